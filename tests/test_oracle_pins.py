"""Pins the CPU oracle: every oracle function is checked against the reference's own
test vectors (transcribed into tests/golden/reference_tests.json), the data fixtures
the reference ships (blobs.tif, mini_blobs + CLIJ CSV, BBBC022 crop), and
independent mathematical checks (closed forms, brute force, scipy partitions)."""
import numpy as np
import pytest
import scipy.ndimage as ndi

from conftest import load_npz
from oracle import cle_oracle as O

A = np.asarray


# ----------------------------------------------------------------- Gaussian
def test_blur_delta_closed_form(ref_tests):
    t = ref_tests["gaussian_blur_delta"]
    out = O.gaussian_blur(A(t["input"], np.float32), *t["sigma"])
    assert out.min() > t["min_gt"]
    assert t["max_gt"] < out.max() < t["max_lt"]
    # centre weight of a normalised 9-tap Gaussian (sigma=1), squared (two passes)
    k = np.arange(-4, 5)
    w0 = 1.0 / np.exp(-k * k / 2.0).sum()
    assert abs(out.max() - 100 * w0 * w0) < 1e-4


def test_difference_of_gaussian_bounds(ref_tests):
    t = ref_tests["difference_of_gaussian"]
    a = A(t["input"], np.float32)
    d = O.gaussian_blur(a, *t["sigma1"]) - O.gaussian_blur(a, *t["sigma2"])
    assert t["min_gt"] < d.min() < t["min_lt"]
    assert t["max_gt"] < d.max() < t["max_lt"]


def _brute_blur(a, sig):
    """Direct (non-separable) convolution in float64 with the clamped product kernel."""
    a3 = a[None] if a.ndim == 2 else a
    D, H, W = a3.shape
    ws = []
    for s in sig:
        if s > 0:
            n = O.sigma_to_kernel_size(s)
            c = (n - 1) // 2
            v = np.arange(-c, c + 1)
            h = np.exp(-(v * v) / (2.0 * s * s))
            ws.append((v, h / h.sum()))
        else:
            ws.append((np.array([0]), np.array([1.0])))
    if a.ndim == 2:
        ws[2] = (np.array([0]), np.array([1.0]))
    out = np.zeros(a3.shape)
    for vz, hz in zip(*ws[2]):
        for vy, hy in zip(*ws[1]):
            for vx, hx in zip(*ws[0]):
                zi = np.clip(np.arange(D) + vz, 0, D - 1)
                yi = np.clip(np.arange(H) + vy, 0, H - 1)
                xi = np.clip(np.arange(W) + vx, 0, W - 1)
                out += hz * hy * hx * a3[np.ix_(zi, yi, xi)]
    return out[0] if a.ndim == 2 else out


@pytest.mark.parametrize("shape,sig", [((9, 13), (1.0, 1.5, 0)), ((5, 7, 6), (0.7, 1.2, 0.9)), ((4, 6, 5), (1.0, 0, 2.0))])
def test_blur_vs_bruteforce(shape, sig):
    rng = np.random.default_rng(1)
    a = rng.random(shape).astype(np.float32) * 100
    out = O.gaussian_blur(a, *sig)
    ref = _brute_blur(a.astype(np.float64), sig)
    np.testing.assert_allclose(out, ref, rtol=1e-5, atol=1e-4)


def test_blur_constant_invariant():
    a = np.full((6, 7, 8), 3.25, np.float32)
    np.testing.assert_allclose(O.gaussian_blur(a, 2, 2, 2), a, rtol=1e-6)


def test_sigma_to_kernel_size():
    assert [O.sigma_to_kernel_size(s) for s in (0, 0.5, 1, 1.5, 2, 3.5, 5)] == [1, 5, 9, 13, 17, 29, 41]


# ----------------------------------------------------------------- mean_box
def test_mean_box_reference(ref_tests):
    t = ref_tests["mean_box"]
    out = O.mean_box(A(t["input"], np.float32), *t["radius"])
    assert np.allclose(out, A(t["expected"]))


def _brute_mean_box(a, rad):
    """Window mean with literal loops in float64: per axis the window of 2r+1 reads
    clamped to the image (the separable kernels' clamp-to-edge reads)."""
    a3 = a[None] if a.ndim == 2 else a
    D, H, W = a3.shape
    rx, ry, rz = rad
    if a.ndim == 2:
        rz = 0
    out = np.zeros(a3.shape)
    for z in range(D):
        for y in range(H):
            for x in range(W):
                s, n = 0.0, 0
                for dz in range(-rz, rz + 1):
                    for dy in range(-ry, ry + 1):
                        for dx in range(-rx, rx + 1):
                            s += float(a3[min(max(z + dz, 0), D - 1), min(max(y + dy, 0), H - 1),
                                          min(max(x + dx, 0), W - 1)])
                            n += 1
                out[z, y, x] = s / n
    return out[0] if a.ndim == 2 else out


@pytest.mark.parametrize("shape,rad", [((9, 11), (1, 2, 0)), ((5, 7, 9), (2, 1, 1)), ((4, 6, 3), (0, 3, 2)),
                                       ((7, 5), (3, 0, 1)), ((3, 4, 8), (1, 1, 1))])
def test_mean_box_vs_float64_bruteforce(shape, rad):
    # random values up to the border: the clamp-to-edge window, the 2r+1 divisor and the
    # radius-0 axes are all visible
    rng = np.random.default_rng(sum(shape) + sum(rad))
    a = (rng.random(shape) * 100).astype(np.float32)
    np.testing.assert_allclose(O.mean_box(a, *rad), _brute_mean_box(a, rad), rtol=1e-5, atol=1e-5)


# ----------------------------------------------------------------- maxima
def test_detect_maxima_reference(ref_tests):
    t = ref_tests["detect_maxima_box"]
    out = O.detect_maxima_box(A(t["input"], np.float32), 0, 0, 0, np.float32)
    assert np.array_equal(out, A(t["expected"], np.float32))


def _brute_maxima(a):
    """Straight transcription of the loop as nested Python loops."""
    a3 = a[None] if a.ndim == 2 else a
    D, H, W = a3.shape
    zr = (-1, 0, 1) if a.ndim == 3 else (0,)
    out = np.zeros(a3.shape, np.uint8)
    for z in range(D):
        for y in range(H):
            for x in range(W):
                lm = np.float32(a3[z, y, x] - np.float32(1))
                pos = (z, y, x)
                for dx in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dz in zr:
                            q = (z + dz, y + dy, x + dx)
                            if min(q) < 0:
                                continue
                            v = a3[min(q[0], D - 1), min(q[1], H - 1), min(q[2], W - 1)]
                            if v > lm:
                                lm, pos = v, q
                out[z, y, x] = pos == (z, y, x)
    return out[0] if a.ndim == 2 else out


@pytest.mark.parametrize("shape", [(7, 9), (4, 5, 6)])
def test_detect_maxima_bruteforce_plateaus(shape):
    rng = np.random.default_rng(2)
    a = rng.integers(0, 4, shape).astype(np.float32)  # many ties / plateaus
    assert np.array_equal(O.detect_maxima_box(a), _brute_maxima(a))


# ----------------------------------------------------------------- histogram / Otsu
@pytest.mark.parametrize("key", ["histogram_2d", "histogram_3d", "histogram_3d_2"])
def test_histogram_reference(ref_tests, key):
    t = ref_tests[key]
    h = O.histogram(A(t["input"], np.float32), t["num_bins"])
    assert np.allclose(h, t["expected"])


def test_histogram_counts_and_integer_bins():
    rng = np.random.default_rng(3)
    a = rng.integers(0, 256, (5, 17, 19)).astype(np.float32)
    a.flat[0], a.flat[1] = 0, 255
    h = O.histogram(a, 256)
    assert h.sum() == a.size
    # for integer data spanning exactly [0,255] the bin of v is v itself
    assert np.array_equal(h, np.bincount(a.astype(int).ravel(), minlength=256))


def _brute_otsu(hist, mn, mx):
    bc = np.arange(256) * np.float64(np.float32(mx) - np.float32(mn)) / 255 + np.float64(mn)
    best, bi = -1.0, 0
    for i in range(255):
        w1, w2 = hist[: i + 1].sum(), hist[i + 1:].sum()
        if w1 == 0 or w2 == 0:
            continue
        m1 = (hist[: i + 1] * bc[: i + 1]).sum() / w1
        m2 = (hist[i + 1:] * bc[i + 1:]).sum() / w2
        v = w1 * w2 * (m1 - m2) ** 2
        if v > best * (1 + 1e-12):
            best, bi = v, i
    return bc[bi]


def test_otsu_vs_bruteforce():
    rng = np.random.default_rng(4)
    a = np.concatenate([rng.normal(50, 8, 3000), rng.normal(160, 15, 1500)]).astype(np.float32)
    mn, mx = a.min(), a.max()
    h = O.histogram(a.reshape(1, -1), 256, mn, mx, False).astype(np.float64)
    t = O.otsu_threshold_from_histogram(h, mn, mx)
    assert abs(t - _brute_otsu(h, mn, mx)) < 1e-9


def test_otsu_constant_image():
    a = np.full((4, 5), 7.0, np.float32)
    out, t = O.threshold_otsu(a, return_threshold=True)
    assert t == 7.0 and out.sum() == 0


# ----------------------------------------------------------------- elementwise
def test_elementwise_reference(ref_tests):
    t = ref_tests["greater_constant"]
    assert np.array_equal(O.greater_constant(A(t["input"]), t["constant"]), A(t["expected"]))
    t = ref_tests["binary_and"]
    assert np.array_equal(O.binary_and(A(t["a"]), A(t["b"])), A(t["expected"]))
    t = ref_tests["mask"]
    assert np.array_equal(O.mask(A(t["input"]), A(t["mask"])), A(t["expected"]))
    t = ref_tests["convert_sat_uint8"]
    assert np.array_equal(O.convert_sat(A(t["input"], np.float32), np.uint8), A(t["expected"]))


# ----------------------------------------------------------------- CCL / relabel
def test_pixel_index_reference(ref_tests):
    t = ref_tests["set_nonzero_pixels_to_pixel_index"]
    assert np.array_equal(O.set_nonzero_pixels_to_pixel_index(A(t["input"])), A(t["expected"]))


@pytest.mark.parametrize("kind", ["box", "diamond"])
def test_nonzero_minimum_reference(ref_tests, kind):
    t = ref_tests["nonzero_minimum_" + kind]
    a = A(t["input"], np.float32)
    out, _ = O.nonzero_minimum(a, np.zeros_like(a), kind)
    assert np.allclose(out, A(t["expected"]), atol=1e-5)


@pytest.mark.parametrize("kind", ["box", "diamond"])
def test_ccl_reference(ref_tests, kind):
    t = ref_tests["connected_components_labeling_" + kind]
    out = O.connected_components_labeling(A(t["input"]), kind)
    assert np.array_equal(out, A(t["expected"]))


def test_ccl_blobs_62(ref_tests):
    img = load_npz("blobs.npz")["image"]
    blurred = O.gaussian_blur(img.astype(np.float32), 1, 1)
    binary = O.threshold_otsu(blurred)
    labels = O.connected_components_labeling(binary)
    assert labels.max() == ref_tests["connected_components_labeling_box_blobs"]["num_labels"]


def _check_partition_and_order(binary, labels, kind):
    ndim = binary.ndim
    structure = np.ones((3,) * ndim) if kind == "box" else ndi.generate_binary_structure(ndim, 1)
    sp, n = ndi.label(binary != 0, structure=structure)
    assert labels.max() == n
    # same partition: a bijection between scipy ids and our ids
    pairs = np.unique(np.stack([sp[binary != 0], labels[binary != 0].astype(np.int64)]), axis=1)
    assert pairs.shape[1] == n and len(np.unique(pairs[0])) == n and len(np.unique(pairs[1])) == n
    # order: ids ascend with each component's minimum x-major key
    keys = O.pixel_keys(binary)
    mins = ndi.minimum(keys, labels, index=np.arange(1, n + 1))
    assert np.all(np.diff(mins) > 0)


@pytest.mark.parametrize("shape,kind", [((23, 31), "box"), ((9, 11, 13), "box"), ((19, 17), "diamond"), ((7, 8, 9), "diamond")])
def test_ccl_partition_vs_scipy(shape, kind):
    rng = np.random.default_rng(5)
    b = (rng.random(shape) < 0.35).astype(np.uint8)
    labels = O.connected_components_labeling(b, kind)
    _check_partition_and_order(b, labels, kind)


def test_relabel_reference(ref_tests):
    t = ref_tests["relabel_sequential"]
    assert np.array_equal(O.relabel_sequential(A(t["input"])), A(t["expected"]))


# ----------------------------------------------------------------- Voronoi
@pytest.mark.parametrize("kind", ["box", "diamond"])
def test_onlyzero_reference(ref_tests, kind):
    t = ref_tests["onlyzero_overwrite_maximum_" + kind]
    out, _ = O.onlyzero_overwrite_maximum(A(t["input"], np.float32), kind)
    assert np.array_equal(out, A(t["expected"]))


@pytest.mark.parametrize("key", ["dilate_labels_2d", "dilate_labels_3d"])
def test_dilate_labels_reference(ref_tests, key):
    # radius 1 = one box step: the 3D case fixes the 27-voxel box neighbourhood across z
    t = ref_tests[key]
    out, changed = O.onlyzero_overwrite_maximum(A(t["input"], np.float32), "box")
    assert changed and np.array_equal(out, A(t["expected"]))
    assert np.array_equal(O.dilate_labels(A(t["input"]), radius=1), A(t["expected"]))


def test_dilate_labels_radius_bounds_growth():
    # radius r grows a single seed to the alternating box/diamond ball of r steps; r
    # beyond convergence changes nothing more
    a = np.zeros((5, 9, 9), np.uint32)
    a[2, 4, 4] = 7
    one = O.dilate_labels(a, 1)
    assert (one == 7).sum() == 27
    two = O.dilate_labels(a, 2)
    # box then diamond: the 3x3x3 cube plus its 6 face layers of 3x3 voxels
    assert (two == 7).sum() == 27 + 6 * 9
    assert np.array_equal(O.dilate_labels(a, 50), O.extend_labeling_via_voronoi(a))


def test_masked_voronoi_reference(ref_tests):
    t = ref_tests["masked_voronoi_labeling"]
    out = O.masked_voronoi_labeling(A(t["input"]), A(t["mask"]))
    assert np.array_equal(out, A(t["expected"]))


def test_voronoi_labeling_reference(ref_tests):
    # voronoi_labeling = masked Voronoi with an all-ones mask
    t = ref_tests["voronoi_labeling"]
    b = A(t["input"])
    out = O.masked_voronoi_labeling(b, np.ones_like(b))
    assert np.array_equal(out, A(t["expected"]))


def test_extend_labeling_via_voronoi_reference(ref_tests):
    t = ref_tests["extend_labeling_via_voronoi"]
    out = O.extend_labeling_via_voronoi(A(t["input"]))
    assert out.dtype == np.uint32 and np.array_equal(out, A(t["expected"]))


def test_voronoi_labeling_via_extend(ref_tests):
    # the reference composes voronoi_labeling as CCL box + extend_labeling_via_voronoi
    t = ref_tests["voronoi_labeling"]
    assert np.array_equal(O.voronoi_labeling(A(t["input"])), A(t["expected"]))


def test_extend_labeling_is_nearest_label_in_chebyshev_octagon():
    # the growth never changes a non-zero voxel, fills every zero voxel that a
    # label can reach, and is idempotent on its own output
    rng = np.random.default_rng(7)
    for shape in ((9, 13), (5, 7, 6)):
        lab = np.zeros(shape, np.uint32)
        idx = rng.choice(lab.size, 4, replace=False)
        lab.flat[idx] = [3, 9, 1, 20]
        out = O.extend_labeling_via_voronoi(lab)
        assert np.array_equal(out.flat[idx], lab.flat[idx])
        assert (out > 0).all()
        assert np.array_equal(O.extend_labeling_via_voronoi(out), out)
        # a single seed fills the whole image with its value
        one = np.zeros(shape, np.uint32)
        one.flat[idx[0]] = 5
        assert (O.extend_labeling_via_voronoi(one) == 5).all()


def test_voronoi_otsu_reference(ref_tests):
    t = ref_tests["voronoi_otsu_labeling"]
    out = O.voronoi_otsu_labeling(A(t["input"]), t["spot_sigma"], t["outline_sigma"])
    assert np.array_equal(out, A(t["expected"]))


def test_gauss_otsu_reference(ref_tests):
    t = ref_tests["gauss_otsu_labeling"]
    out = O.gauss_otsu_labeling(A(t["input"]), t["outline_sigma"])
    assert np.array_equal(out, A(t["expected"]))


def test_bbbc022_spot_counts(ref_tests):
    t = ref_tests["bbbc022_spots"]
    img = load_npz("bbbc022_crop.npz")["image"].astype(np.float32)
    blurred = O.gaussian_blur(img, t["spot_sigma"], t["spot_sigma"], t["spot_sigma"])
    spots = O.detect_maxima_box(blurred)
    assert spots.sum() == t["detected_spots"]
    binary = O.threshold_otsu(O.gaussian_blur(img, t["outline_sigma"], t["outline_sigma"], t["outline_sigma"]))
    assert O.binary_and(binary, spots).sum() == t["selected_spots"]


# ----------------------------------------------------------------- label_spots
@pytest.mark.parametrize("key", ["label_spots_2d", "label_spots_3d"])
def test_label_spots_reference(ref_tests, key):
    t = ref_tests[key]
    assert np.array_equal(O.label_spots(A(t["input"])), A(t["expected"]))


# ----------------------------------------------------------------- statistics
@pytest.mark.parametrize("key", ["statistics_2d", "statistics_3d"])
def test_statistics_reference_tables(ref_tests, key):
    t = ref_tests[key]
    r = O.statistics_of_labelled_pixels(A(t["intensity"]), A(t["labels"]))
    assert list(r.keys()) == O.STAT_KEYS
    for k, v in t["expected"].items():
        assert np.allclose(r[k], v, t["rtol"]), k


def test_statistics_std(ref_tests):
    t = ref_tests["statistics_std"]
    r = O.statistics_of_labelled_pixels(A(t["intensity"]), A(t["labels"]))
    np.testing.assert_almost_equal(r["standard_deviation_intensity"][0], t["std"], decimal=t["decimal"])


def test_statistics_clij_csv(ref_tests):
    t = ref_tests["statistics_clij_csv"]
    fx = load_npz("mini_blobs.npz")
    r = O.statistics_of_background_and_labelled_pixels(fx["image"], fx["labels"])
    table = fx["clij_table"]
    for i, k in enumerate(t["columns"]):
        assert np.allclose(table[:, i], r[k], t["rtol"]), k


def test_statistics_vs_float64_bruteforce():
    rng = np.random.default_rng(6)
    lab = rng.integers(0, 6, (3, 11, 13))
    img = rng.random(lab.shape).astype(np.float32) * 50
    r = O.statistics_of_labelled_pixels(img, lab)
    zz, yy, xx = np.indices(lab.shape)
    for i, l in enumerate(r["label"]):
        m = lab == l
        v = img[m].astype(np.float64)
        assert r["area"][i] == m.sum()
        assert np.isclose(r["sum_intensity"][i], v.sum(), rtol=1e-6)
        assert np.isclose(r["sum_intensity_times_x"][i], (xx[m] * v).sum(), rtol=1e-6)
        assert r["bbox_min_z"][i] == zz[m].min() and r["bbox_max_y"][i] == yy[m].max()
        cx, cy, cz = xx[m].mean(), yy[m].mean(), zz[m].mean()
        d = np.sqrt((xx[m] - cx) ** 2 + (yy[m] - cy) ** 2 + (zz[m] - cz) ** 2)
        assert np.isclose(r["sum_distance_to_centroid"][i], d.sum(), rtol=1e-5)
        assert np.isclose(r["max_distance_to_centroid"][i], d.max(), rtol=1e-5)
        assert np.isclose(r["standard_deviation_intensity"][i], v.std(), rtol=1e-4)
