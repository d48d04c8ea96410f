"""The Voronoi growth's two methods against the oracle and each other: the image-wide
pairs of steps (default) and the asynchronous tile solves.  Cases: masks and seeds that
cross tile borders in every direction, the reference's stop rule (a growth that could
only continue diagonally after an image-wide step without change), growths beyond the
tile solves' step range (the fallback), and the iteration count."""
import ctypes

import numpy as np
import pytest

from oracle import cle_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture
def methods(cle):
    yield ("tiles", "steps")
    cle.set_growth_method("steps")


def _masked(cle, seeds, mask, method):
    from pyclesperanto_prototype_amd import _lib
    cle.set_growth_method(method)
    try:
        b, m = cle.push(seeds), cle.push(mask)
        out = cle.create(b.shape, np.uint32)
        it = ctypes.c_int(-1)
        _lib.call("cle_masked_voronoi_labeling", b.buf(), m.buf(), out.buf(), ctypes.byref(it), cle._ops._S())
        return cle.pull(out), it.value
    finally:
        cle.set_growth_method("steps")


# tile shapes: 62 x 6 x 6 (3D), 62 x 62 (2D): sizes around and across their borders
@pytest.mark.parametrize("shape,p,q", [((70, 130), 0.8, 0.004), ((125, 63), 0.6, 0.01), ((13, 14, 130), 0.7, 0.004),
                                       ((7, 19, 64), 0.9, 0.002), ((25, 12, 61), 0.5, 0.02), ((1, 200, 190), 0.8, 0.003),
                                       ((3, 3, 300), 0.9, 0.005), ((40, 30, 30), 0.95, 0.0005)])
def test_masked_growth_methods_vs_oracle(cle, methods, shape, p, q):
    rng = np.random.default_rng(abs(hash(shape)) % 2 ** 32)
    mask = (rng.random(shape) < p).astype(np.uint8)
    seeds = ((rng.random(shape) < q) & (mask > 0)).astype(np.uint8)
    want, wit = O.masked_voronoi_labeling(seeds, mask, return_iterations=True)
    for m in methods:
        got, it = _masked(cle, seeds, mask, m)
        assert np.array_equal(got, want), (m, int(np.sum(got != want)))
        assert it == wit, (m, it, wit)


@pytest.mark.parametrize("ndim", [2, 3])
def test_stop_rule_diagonal_continuation(cle, methods, ndim):
    # the seeds' region touches a second region only at a corner: after the first
    # region is filled, a diamond step changes nothing anywhere, so the reference stops
    # and the second region stays 0, although a box step would have entered it.  The
    # regions straddle tile borders (x 62, y 6 / 62, z 6).
    if ndim == 2:
        mask = np.zeros((80, 140), np.uint8)
        mask[50:63, 55:63] = 1
        mask[63:75, 63:130] = 1   # touches (62, 62) only diagonally
        seeds = np.zeros_like(mask)
        seeds[50, 55] = 1
    else:
        mask = np.zeros((14, 16, 140), np.uint8)
        mask[2:6, 3:6, 55:62] = 1
        mask[6:12, 6:14, 62:120] = 1
        seeds = np.zeros_like(mask)
        seeds[3, 4, 58] = 1
    want, wit = O.masked_voronoi_labeling(seeds, mask, return_iterations=True)
    second = (mask > 0) & (want == 0)
    assert second.sum() > 100  # the stop rule matters here
    for m in methods:
        got, it = _masked(cle, seeds, mask, m)
        assert np.array_equal(got, want), m
        assert it == wit, (m, it, wit)
    # with a second seed elsewhere that keeps growing, the corner is crossed
    seeds2 = seeds.copy()
    mask2 = mask.copy()
    if ndim == 2:
        mask2[0, 0:40] = 1
        seeds2[0, 0] = 1
    else:
        mask2[0, 0, 0:40] = 1
        seeds2[0, 0, 0] = 1
    want2, wit2 = O.masked_voronoi_labeling(seeds2, mask2, return_iterations=True)
    for m in methods:
        got2, it2 = _masked(cle, seeds2, mask2, m)
        assert np.array_equal(got2, want2), m
        assert it2 == wit2, (m, it2, wit2)


@pytest.mark.parametrize("shape", [(1, 700), (2, 3, 600)])
def test_long_growth_beyond_tile_step_range(cle, methods, shape):
    # one seed at the end of a line: 699 / 599 growth steps, beyond the tile solves'
    # 254-step range, so the call falls back to the pairs of steps
    mask = np.ones(shape, np.uint8)
    seeds = np.zeros(shape, np.uint8)
    seeds[(0,) * (len(shape) - 1) + (0,)] = 1
    seeds[(0,) * (len(shape) - 1) + (5,)] = 1
    want, wit = O.masked_voronoi_labeling(seeds, mask, return_iterations=True)
    assert wit > 254
    for m in methods:
        got, it = _masked(cle, seeds, mask, m)
        assert np.array_equal(got, want), m
        assert it == wit, (m, it, wit)


@pytest.mark.parametrize("shape", [(64, 200), (9, 20, 150)])
def test_extend_labeling_methods(cle, methods, shape):
    rng = np.random.default_rng(7)
    lab = np.zeros(shape, np.float32)
    idx = rng.choice(lab.size, lab.size // 300, replace=False)
    lab.flat[idx] = rng.integers(1, 2 ** 20, idx.size)
    lab.flat[idx[:5]] = -3.0   # negative values never change and never spread
    want = O.extend_labeling_via_voronoi(lab, np.float32)
    for m in methods:
        cle.set_growth_method(m)
        try:
            got = cle.pull(cle.extend_labeling_via_voronoi(lab, cle.create(shape, np.float32)))
        finally:
            cle.set_growth_method("steps")
        assert np.array_equal(got, want), m


def test_batch_methods_identical(cle, methods):
    import torch
    from pyclesperanto_prototype_amd import synthetic
    imgs_t, ss, so, _ = synthetic.mixed_batch(24, seed=91)
    imgs = [cle.asarray(t) for t in imgs_t]
    outs = {}
    for m in methods:
        cle.set_growth_method(m)
        try:
            outs[m], outs[m + "_steps"] = cle.voronoi_otsu_labeling_batch(imgs, ss, so, return_steps=True)
        finally:
            cle.set_growth_method("steps")
    assert outs["tiles_steps"] == outs["steps_steps"]
    for a, b in zip(outs["tiles"], outs["steps"]):
        assert torch.equal(a.tensor, b.tensor)
