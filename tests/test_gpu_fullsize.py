"""The configured workloads at their full sizes, checked through properties that do not
depend on size (the oracle runs on slabs, scipy on the whole binary volume, the rest
on the device).

* 256 x 1024 x 1024 uint16 nuclei stack (5000 nuclei, seed 2): gaussian_blur(2, 2, 2)
  -> threshold_otsu -> connected_components_labeling_box (int32) -> statistics.
* 512 x 2048 x 2048 uint16 stack (100 k nuclei, semi-axes 3-8, 15 % touching pairs,
  seed 3): voronoi_otsu_labeling(2.0, 1.5).
"""
import sys

import numpy as np
import pytest

from oracle import cle_oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def test_plain3d_pipeline_full_size(cle):
    import torch
    import scipy.ndimage as ndi
    from pyclesperanto_prototype_amd import synthetic
    from test_gpu_parity import _ccl_expected_fp32_keys

    shape = (256, 1024, 1024)
    vol_t, info = synthetic.nuclei_3d(shape, 5000, seed=2)
    vol = cle.asarray(vol_t)
    host = vol_t.view(torch.int16).cpu().numpy().view(np.uint16)
    _say(f"full size 3D: {info['nuclei']} nuclei")

    # blur: a slab of 8 slices against the oracle run on those slices + the 8-slice halo
    blurred = cle.gaussian_blur(vol, cle.create(shape, np.float32), 2, 2, 2)
    b = cle.pull(blurred)
    ref = O.gaussian_blur(host[112:136].astype(np.float32), 2, 2, 2)[8:16]
    err = np.abs(b[120:128].astype(np.float64) - ref).max()
    assert err <= 1e-4 * max(1.0, float(np.abs(ref).max())), err
    _say(f"blur slab ok (max abs err {err:.3g})")

    # histogram (per-row partials summed in y order, float32) and the Otsu threshold
    hist = cle.pull(cle.histogram(blurred, num_bins=256))
    want_hist = O.histogram(b, 256)
    assert np.array_equal(hist, want_hist)
    thr = O.otsu_threshold_from_histogram(want_hist, b.min(), b.max())
    binary = cle.threshold_otsu(blurred)
    bh = cle.pull(binary)
    assert np.array_equal(bh, (b > np.float32(thr)).astype(np.uint8))
    _say(f"histogram / Otsu ok (threshold {thr:.3f}, {int(bh.sum())} foreground voxels)")
    del b

    # CCL: scipy components, IDs by the float32-rounded minimum x-major key
    labels = cle.create(shape, np.int32)
    cle.connected_components_labeling_box(binary, labels)
    lab = cle.pull(labels)
    want, n, n_ids = _ccl_expected_fp32_keys(bh)
    assert int(lab.max()) == n_ids
    assert np.array_equal(lab.astype(np.uint32), want)
    _say(f"CCL ok ({n} components, {n_ids} labels)")

    # statistics: areas exact, intensity sums within float32 accumulation of the float64 sum
    stats = cle.statistics_of_labelled_pixels(vol, labels)
    area = np.bincount(lab.ravel(), minlength=n_ids + 1)[1:]
    assert np.array_equal(stats["area"], area.astype(np.float32))
    sums = np.bincount(lab.ravel(), weights=host.ravel().astype(np.float64), minlength=n_ids + 1)[1:]
    np.testing.assert_allclose(stats["sum_intensity"], sums, rtol=1e-4)
    com = ndi.center_of_mass(np.ones(shape, np.uint8), lab, index=[1, n_ids])
    np.testing.assert_allclose(stats["centroid_z"][[0, -1]], [c[0] for c in com], rtol=1e-4, atol=1e-3)
    np.testing.assert_allclose(stats["centroid_x"][[0, -1]], [c[2] for c in com], rtol=1e-4, atol=1e-3)
    _say("statistics ok")


def test_voronoi_otsu_labeling_large_full_size(cle):
    import torch
    from pyclesperanto_prototype_amd import synthetic

    shape = (512, 2048, 2048)
    vol_t, info = synthetic.nuclei_3d(shape, 100_000, semi=(3.0, 8.0), seed=3, touching_fraction=0.15)
    vol = cle.asarray(vol_t)
    _say(f"full size large: {info['nuclei']} nuclei")
    labels = cle.voronoi_otsu_labeling(vol, spot_sigma=2.0, outline_sigma=1.5)
    lab = labels.tensor.view(torch.int32)

    # the stages recomputed on the device: segmentation and seeds
    seg = cle.threshold_otsu(cle.gaussian_blur(vol, sigma_x=1.5, sigma_y=1.5, sigma_z=1.5))
    spots = cle.detect_maxima_box(cle.gaussian_blur(vol, sigma_x=2, sigma_y=2, sigma_z=2))
    seeds = cle.binary_and(spots, seg)
    seed_ids = cle.connected_components_labeling_box(seeds).tensor.view(torch.int32)
    n_seed = int(seed_ids.max())
    _say(f"{n_seed} seed components, {int(lab.max())} labels")
    # labels are the seed components' CCL labels, kept on the seeds, grown only inside
    # the segmentation (checked in slabs: torch indexing stops short of 2^31 elements)
    assert int(lab.max()) == n_seed
    assert int(lab.min()) >= 0
    for z in range(0, shape[0], 64):
        sl = slice(z, z + 64)
        s = seeds.tensor[sl] != 0
        assert torch.equal(lab[sl][s], seed_ids[sl][s]), z
        assert not bool(((lab[sl] != 0) & (seg.tensor[sl] == 0)).any()), z
    # deterministic
    again = cle.voronoi_otsu_labeling(vol, spot_sigma=2.0, outline_sigma=1.5)
    assert torch.equal(again.tensor, labels.tensor)
