"""The size-dependent paths against the oracle, bit for bit.

* Single images of 4096+ Voronoi growth tiles (2048^2 2D, 64x256x256 3D, and a
  256x1024x1024 stack) grow over runs of 16 tiles per block: masked_voronoi_labeling
  and voronoi_otsu_labeling against the C oracle given the same seeds and mask, and
  the device blurs.
* Every member of a mixed batch (2D and 3D, 4096+-tile members included) and every
  request served in a mix (one 256x1024x1024 stack among them) against the oracle's
  stages fed with that image's device blurs; the blurs themselves against the
  oracle's blur within the blur tolerance.
* The full 256x1024x1024 plain method: binary, int32 labels and the 36-column
  statistics table against the C oracle's Otsu, CCL and statistics of the device
  blur.

The numpy oracle checks images up to 200 k voxels, the C restatement (reference
structure: Jacobi growth to a fixed point, per-row statistics) the larger ones.
"""
import sys

import numpy as np
import pytest

from oracle import c_oracle as C
from oracle import cle_oracle as O

pytestmark = pytest.mark.gpu

BLUR_TOL = 1e-4
STATS_RTOL = 1e-5
NUMPY_MAX_VOXELS = 200_000


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _host(t):
    """A device tensor (uint16 or float32) as a float32 numpy array."""
    import torch
    if t.dtype == torch.uint16:
        return t.view(torch.int16).cpu().numpy().view(np.uint16).astype(np.float32)
    return t.float().cpu().numpy()


def oracle_labels_from_blurs(gbs, gbo):
    """voronoi_otsu_labeling downstream of its two blurs (_tier9/_voronoi_otsu_labeling.py:
    59-73): maxima of the spot blur AND the Otsu segmentation of the outline blur,
    masked Voronoi of the seeds inside the segmentation.  Returns (labels, seeds, seg)."""
    if gbs.size <= NUMPY_MAX_VOXELS:
        spots = O.detect_maxima_box(gbs, 0, 0, 0, np.uint8)
        seg = O.threshold_otsu(gbo, np.uint8)
        seeds = O.binary_and(spots, seg, np.uint8)
        return O.masked_voronoi_labeling(seeds, seg), seeds, seg
    spots = C.detect_maxima(gbs)
    seg, _ = C.threshold_otsu(gbo)
    seeds = (spots & seg).astype(np.uint8)
    lab, _ = C.masked_voronoi(seeds, seg)
    return lab, seeds, seg


def _check_blur(got, src, sigma):
    """The device blur against the oracle's (the C restatement for large images, on a
    slab of at most 24 slices + the kernel halo for stacks)."""
    want_full = None
    if got.ndim == 3 and got.shape[0] > 24:
        c = (O.sigma_to_kernel_size(sigma) - 1) // 2
        z0 = got.shape[0] // 2 - 12
        lo, hi = max(0, z0 - c), min(got.shape[0], z0 + 24 + c)
        ref = C.gaussian_blur(src[lo:hi], sigma, sigma, sigma)[z0 - lo:z0 - lo + 24]
        got = got[z0:z0 + 24]
    else:
        want_full = O.gaussian_blur(src, sigma, sigma, sigma) if src.size <= NUMPY_MAX_VOXELS else \
            C.gaussian_blur(src, sigma, sigma, sigma)
        ref = want_full
    err = float(np.abs(got.astype(np.float64) - ref).max()) if got.size else 0.0
    assert err <= BLUR_TOL * max(1.0, float(np.abs(ref).max())), (got.shape, sigma, err)


def _device_blurs(cle, img, ss, so):
    gbs = cle.pull(cle.gaussian_blur(img, None, ss, ss, ss))
    gbo = cle.pull(cle.gaussian_blur(img, None, so, so, so))
    return gbs, gbo


def _tiles(shape):
    if len(shape) == 3 and shape[0] > 1:
        return -(-shape[0] // 8) * -(-shape[1] // 8) * -(-shape[2] // 16)
    return -(-shape[-2] // 32) * -(-shape[-1] // 32)


def _nuclei(shape, seed):
    from pyclesperanto_prototype_amd import synthetic
    if len(shape) == 2:
        return synthetic.nuclei_overlapping(shape, 2e-4, seed=seed)[0]
    n = int(np.prod(shape) / 50_000)
    return synthetic.nuclei_3d(shape, n, seed=seed)[0]


# ----------------------------------------------------------------- single images, raster mode
@pytest.mark.parametrize("shape", [(2048, 2048), (64, 256, 256), (256, 1024, 1024)])
def test_single_image_raster_growth_vs_oracle(cle, shape):
    import torch
    assert _tiles(shape) >= 4096
    vol_t = _nuclei(shape, 70 + len(shape))
    img = cle.asarray(vol_t)
    ss, so = 2.0, 1.5
    _say(f"{shape}: {_tiles(shape)} growth tiles")
    gbs, gbo = _device_blurs(cle, img, ss, so)
    want, seeds, seg = oracle_labels_from_blurs(gbs, gbo)
    assert seeds.sum() > 10 and want.max() > 10
    # masked_voronoi_labeling entry point on the oracle's own seeds and mask
    got_mv = cle.pull(cle.masked_voronoi_labeling(seeds, seg))
    assert np.array_equal(got_mv, want), int(np.sum(got_mv != want))
    # the whole pipeline: the device's own maxima / Otsu / Voronoi from its blurs
    got = cle.pull(cle.voronoi_otsu_labeling(img, None, ss, so))
    assert got.dtype == np.uint32 and np.array_equal(got, want), int(np.sum(got != want))
    _say(f"{shape}: {int(want.max())} labels identical to the oracle")
    src = _host(vol_t)
    _check_blur(gbs, src, ss)
    _check_blur(gbo, src, so)
    del img, vol_t
    torch.cuda.empty_cache()


# ----------------------------------------------------------------- batch members
def test_batch_members_vs_oracle(cle):
    import torch
    from pyclesperanto_prototype_amd import synthetic

    imgs_t, ss, so, info = synthetic.mixed_batch(40, seed=41)
    # members of 4096+ growth tiles, a dense (percolated) stack and a float32 source
    extra = [(synthetic.nuclei_3d((64, 256, 256), 300, seed=42)[0], 2.0, 1.5),
             (synthetic.nuclei_overlapping((2048, 2048), 5e-4, seed=43)[0], 3.0, 1.0),
             (synthetic.nuclei_overlapping((40, 200, 200), 1e-3, seed=44)[0], 1.5, 0.8),
             (synthetic.nuclei_overlapping((300, 260), 6e-4, seed=45)[0].float(), 2.5, 1.2)]
    imgs = [cle.asarray(t) for t in imgs_t] + [cle.asarray(t) for t, _, _ in extra]
    ss = [float(v) for v in ss] + [e[1] for e in extra]
    so = [float(v) for v in so] + [e[2] for e in extra]
    out = cle.voronoi_otsu_labeling_batch(imgs, ss, so)
    _say(f"batch of {len(imgs)} images, {sum(i.size for i in imgs)} voxels")
    n_big = 0
    for k, im in enumerate(imgs):
        gbs, gbo = _device_blurs(cle, im, ss[k], so[k])
        want, _, _ = oracle_labels_from_blurs(gbs, gbo)
        got = cle.pull(out[k])
        assert np.array_equal(got, want), (k, im.shape, ss[k], so[k], int(np.sum(got != want)))
        n_big += _tiles(im.shape) >= 4096
        if k % 8 == 0 or k >= len(imgs) - len(extra):
            src = _host(im.tensor)
            _check_blur(gbs, src, ss[k])
            _check_blur(gbo, src, so[k])
    assert n_big >= 2
    _say(f"all {len(imgs)} members equal the oracle ({n_big} of 4096+ tiles)")
    del out, imgs, imgs_t
    torch.cuda.empty_cache()


# ----------------------------------------------------------------- served requests
def test_served_requests_vs_oracle(cle):
    import torch
    from pyclesperanto_prototype_amd import serving, synthetic

    rng = np.random.default_rng(80)
    small_t, ss, so, _ = synthetic.mixed_batch(12, seed=81)
    reqs = [(cle.asarray(t), float(a), float(b)) for t, a, b in zip(small_t, ss, so)]
    big = cle.asarray(synthetic.nuclei_3d((256, 1024, 1024), 5000, seed=82)[0])
    reqs.insert(5, (big, 2.0, 1.5))
    arrivals = serving.poisson_trace(len(reqs), qps=3000.0, seed=83)
    with serving.SegmentationServer(workers=2) as srv:
        done = serving.replay(srv, reqs, arrivals)
        sizes = [b[0] for b in srv.batches]
    _say(f"served {len(reqs)} requests in batches {sizes}")
    for k, (r, (im, a, b)) in enumerate(zip(done, reqs)):
        gbs, gbo = _device_blurs(cle, im, a, b)
        want, _, _ = oracle_labels_from_blurs(gbs, gbo)
        got = cle.pull(r.result())
        assert np.array_equal(got, want), (k, im.shape, int(np.sum(got != want)))
    _say("every served result equals the oracle")
    del done, reqs, big
    torch.cuda.empty_cache()


# ----------------------------------------------------------------- plain method, full size
def test_plain3d_full_size_labels_and_table_vs_c_oracle(cle):
    import torch
    from pyclesperanto_prototype_amd import synthetic
    from pyclesperanto_prototype_amd._ops import _TABLE_COLUMNS

    shape = (256, 1024, 1024)
    vol_t, _ = synthetic.nuclei_3d(shape, 5000, seed=2)
    vol = cle.asarray(vol_t)
    blurred = cle.gaussian_blur(vol, cle.create(shape, np.float32), 2, 2, 2)
    binary = cle.threshold_otsu(blurred)
    labels = cle.create(shape, np.int32)
    cle.connected_components_labeling_box(binary, labels)
    stats = cle.statistics_of_labelled_pixels(vol, labels)
    b = cle.pull(blurred)
    seg, thr = C.threshold_otsu(b)
    assert cle.otsu_threshold(blurred) == thr
    assert np.array_equal(cle.pull(binary), seg)
    want_lab, _ = C.ccl(seg)
    lab = cle.pull(labels)
    assert np.array_equal(lab.astype(np.uint32), want_lab)
    _say(f"C2 labels identical ({int(lab.max())} labels, threshold {thr:.3f})")
    del b, blurred, binary
    table = C.statistics_table(_host(vol_t), want_lab)
    assert table.shape[1] == len(stats["label"])
    for i, k in enumerate(_TABLE_COLUMNS):
        want = table[i]
        got = np.asarray(stats["label"] if k == "original_label" else stats[k], np.float64)
        np.testing.assert_allclose(got, want, rtol=STATS_RTOL, atol=0, equal_nan=True, err_msg=k)
    _say("C2 statistics table equal to the C oracle within 1e-5")
    del vol, labels, vol_t
    torch.cuda.empty_cache()
