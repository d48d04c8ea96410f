"""Pins the C restatement (CPU baseline / large-size checker) against the numpy oracle."""
import numpy as np
import pytest

from oracle import c_oracle as C
from oracle import cle_oracle as O


def blobs(shape, n, seed):
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    img = np.full(shape, 10.0)
    for _ in range(n):
        c = [rng.uniform(0, s) for s in shape]
        s = rng.uniform(2, 5)
        img += rng.uniform(100, 200) * np.exp(-sum((g - ci) ** 2 for g, ci in zip(grids, c)) / (2 * s * s))
    return np.clip(img + rng.normal(0, 5, shape), 0, 255).astype(np.float32)


@pytest.mark.parametrize("shape", [(40, 50), (7, 20, 25)])
def test_c_blur_maxima_otsu(shape):
    a = blobs(shape, 6, 1)
    b = C.gaussian_blur(a, 1.5, 1.5, 1.5)
    ref = O.gaussian_blur(a, 1.5, 1.5, 1.5)
    assert np.abs(b - ref).max() <= 1e-4 * np.abs(ref).max()
    assert np.array_equal(C.detect_maxima(ref), O.detect_maxima_box(ref))
    out, t = C.threshold_otsu(ref)
    want, tw = O.threshold_otsu(ref, return_threshold=True)
    assert t == tw and np.array_equal(out, want)


@pytest.mark.parametrize("shape,diamond", [((33, 41), False), ((9, 14, 15), False), ((30, 30), True)])
def test_c_ccl(shape, diamond):
    b = (np.random.default_rng(2).random(shape) < 0.4).astype(np.uint8)
    got, _ = C.ccl(b, diamond)
    assert np.array_equal(got, O.connected_components_labeling(b, "diamond" if diamond else "box"))


@pytest.mark.parametrize("shape", [(60, 60), (10, 30, 30)])
def test_c_voronoi_otsu_and_stats(shape):
    a = blobs(shape, 8, 3)
    lab = C.voronoi_otsu_labeling(a, 2.0, 1.0)
    want = O.voronoi_otsu_labeling(a, 2.0, 1.0)
    assert np.array_equal(lab, want)
    t = C.statistics_table(a, lab)
    r = O.statistics_of_labelled_pixels(a, want)
    from pyclesperanto_prototype_amd._ops import _TABLE_COLUMNS
    for i, k in enumerate(_TABLE_COLUMNS):
        if k == "original_label":
            assert np.array_equal(t[i], r["label"])
            continue
        np.testing.assert_allclose(t[i], r[k], rtol=1e-5, equal_nan=True, err_msg=k)
