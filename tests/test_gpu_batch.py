"""Batched voronoi_otsu_labeling: every image of a mixed batch (2D and 3D, several
sizes, source types and sigmas) must equal the single-image device path bit for bit,
and a captured plan's replays must equal the eager batch.  The members against the
oracle are in test_gpu_oracle_large.py (test_batch_members_vs_oracle)."""
import numpy as np
import pytest

from conftest import blobs_image

pytestmark = pytest.mark.gpu


def _mixed_batch(rng, cle):
    shapes = [(96, 128), (64, 64), (300, 257), (17, 33, 65), (12, 40, 70), (1, 50, 60), (200, 200), (9, 64, 64),
              (130, 70), (5, 20, 300)]
    imgs, ss, so = [], [], []
    for k, sh in enumerate(shapes):
        a = blobs_image(sh, 6 + k % 5, 100 + k).astype(np.float32)
        a = (a * 40 + rng.normal(0, 3, sh)).clip(0, 65535)
        if k % 3 == 0:
            imgs.append(cle.asarray(a.astype(np.uint16)))     # native uint16 source
        else:
            imgs.append(cle.push(a))                          # float32, as push gives it
        ss.append(float(rng.uniform(1.0, 5.0)))
        so.append(float(rng.uniform(0.5, 2.0)))
    ss[2] = 4.6   # 37 taps: wider than the isotropic kernel, runs the general path
    return imgs, ss, so


def test_batch_equals_single_image_path(cle):
    rng = np.random.default_rng(50)
    imgs, ss, so = _mixed_batch(rng, cle)
    out, steps = cle.voronoi_otsu_labeling_batch(imgs, ss, so, return_steps=True)
    assert steps > 0
    for k, (im, a, b) in enumerate(zip(imgs, ss, so)):
        want = cle.pull(cle.voronoi_otsu_labeling(im, None, a, b))
        got = cle.pull(out[k])
        assert got.dtype == np.uint32 and got.shape == want.shape
        assert np.array_equal(got, want), (k, im.shape, a, b)
        assert got.max() > 0, k


def test_batch_equals_single_multi_tile_raster(cle):
    # single images of 4096+ growth tiles grow in raster order over runs of 16 tiles per
    # block; a batch of them grows from work lists: both give the same labels
    import torch
    from pyclesperanto_prototype_amd import synthetic
    imgs = [cle.asarray(synthetic.nuclei_3d(sh, n, seed=60 + k)[0])
            for k, (sh, n) in enumerate((((64, 256, 256), 300), ((48, 300, 330), 250)))]
    imgs.append(cle.asarray(synthetic.nuclei_overlapping((2048, 2048), 2e-4, seed=62)[0]))
    ss, so = [2.0, 1.5, 3.0], [1.5, 1.0, 2.0]
    out = cle.voronoi_otsu_labeling_batch(imgs, ss, so)
    for k, im in enumerate(imgs):
        want = cle.voronoi_otsu_labeling(im, None, ss[k], so[k])
        assert torch.equal(out[k].tensor, want.tensor), k
        assert int(want.tensor.view(torch.int32).max()) > 10, k


def test_batch_many_small_images(cle):
    # 4100 small images (a block finds its image among them in three rounds of the
    # 64-ary wavefront search; empty-tile images and 1-voxel rows included)
    import torch
    rng = np.random.default_rng(63)
    imgs, ss, so = [], [], []
    for k in range(4100):
        h, w = int(rng.integers(1, 48)), int(rng.integers(1, 48))
        a = np.zeros((h, w), np.float32)
        for _ in range(int(rng.integers(0, 3))):
            y, x = int(rng.integers(0, h)), int(rng.integers(0, w))
            a[max(0, y - 2):y + 3, max(0, x - 2):x + 3] += float(rng.uniform(50, 200))
        imgs.append(torch.from_numpy(a))
        ss.append(float(rng.uniform(0.5, 2.0)))
        so.append(float(rng.uniform(0.5, 1.5)))
    dev = [cle.asarray(t.cuda()) for t in imgs]
    out = cle.voronoi_otsu_labeling_batch(dev, ss, so)
    for k in range(0, 4100, 7):
        want = cle.voronoi_otsu_labeling(dev[k], None, ss[k], so[k])
        assert torch.equal(out[k].tensor, want.tensor), k


def test_batch_single_member_and_repeat(cle):
    """A batch of one equals the single call; running a batch twice gives the same."""
    rng = np.random.default_rng(51)
    a = cle.push((blobs_image((40, 60, 50), 8, 7) * 30).astype(np.float32))
    one = cle.pull(cle.voronoi_otsu_labeling_batch([a], 2.0, 1.0)[0])
    assert np.array_equal(one, cle.pull(cle.voronoi_otsu_labeling(a, None, 2.0, 1.0)))
    imgs, ss, so = _mixed_batch(rng, cle)
    r1 = [cle.pull(x) for x in cle.voronoi_otsu_labeling_batch(imgs, ss, so)]
    r2 = [cle.pull(x) for x in cle.voronoi_otsu_labeling_batch(imgs, ss, so)]
    assert all(np.array_equal(x, y) for x, y in zip(r1, r2))


def test_batch_large_member_uses_row_histogram(cle):
    """A member of 2^24 voxels takes the per-row histogram of the single-image Otsu."""
    import torch
    from pyclesperanto_prototype_amd import synthetic
    big, _ = synthetic.nuclei_3d((64, 512, 512), 400, seed=52)
    small = cle.asarray((blobs_image((80, 90), 5, 3) * 50).astype(np.uint16))
    imgs = [cle.asarray(big), small]
    out = cle.voronoi_otsu_labeling_batch(imgs, [2.0, 1.5], [1.5, 1.0])
    for k, (a, b) in enumerate([(2.0, 1.5), (1.5, 1.0)]):
        want = cle.voronoi_otsu_labeling(imgs[k], None, a, b)
        assert torch.equal(out[k].tensor, want.tensor), k


def _refill(cle, imgs, shift):
    """New contents in the same device buffers (rolled images)."""
    import torch
    for k, im in enumerate(imgs):
        a = np.roll(cle.pull(im), shift + k, axis=-1)
        t = im.tensor
        if a.dtype == np.uint16:
            t.view(torch.int16).copy_(torch.from_numpy(np.ascontiguousarray(a).view(np.int16)).to(t.device))
        else:
            t.copy_(torch.from_numpy(np.ascontiguousarray(a)).to(t.device))


@pytest.mark.parametrize("graph_pairs,single", [(32, False), (1, False), (2, True)])
def test_batch_plan_replays_equal_eager(cle, graph_pairs, single):
    # the captured graph replayed three times on new source contents, with eager work
    # and synchronisations between the replays; every replay equals the eager batch
    # (graph_pairs = 1 forces the eager tail of the growth after the replay)
    import torch
    rng = np.random.default_rng(51)
    imgs, ss, so = _mixed_batch(rng, cle)
    if single:
        imgs, ss, so = imgs[3:4], ss[3:4], so[3:4]
    plan = cle.voronoi_otsu_labeling_batch_plan(imgs, ss, so, graph_pairs=graph_pairs)
    try:
        want = [cle.pull(o) for o in cle.voronoi_otsu_labeling_batch(imgs, ss, so)]
        for d, w in zip(plan.destinations, want):  # building ran the batch
            assert np.array_equal(cle.pull(d), w)
        for rep in range(3):
            _refill(cle, imgs, rep + 1)
            torch.cuda.synchronize()
            want = [cle.pull(o) for o in cle.voronoi_otsu_labeling_batch(imgs, ss, so)]
            plan.run()
            assert plan.steps > 0
            for k, d in enumerate(plan.destinations):
                assert np.array_equal(cle.pull(d), want[k]), (rep, k, imgs[k].shape)
    finally:
        plan.close()
    with pytest.raises(ValueError):
        plan.run()


def test_batched_blur_launch_shapes_identical(cle):
    # the batched blur under the default launch shapes, while the autotuner times every
    # group's shapes (tile height 32 / 16 rows, register budget, z chunking), and under
    # the tuned shapes gives the same bits, equal to each image's single-image blur
    import torch
    from oracle import cle_oracle as O
    rng = np.random.default_rng(70)
    imgs, sig = [], []
    for k, sh in enumerate([(70, 130), (300, 257), (20, 37, 70), (12, 100, 90), (1, 50, 60), (9, 64, 64),
                            (64, 64), (33, 47), (5, 20, 300), (200, 200)]):
        a = rng.integers(0, 4000, sh)
        if k % 3 == 0:
            imgs.append(cle.asarray(torch.from_numpy(a.astype(np.int16)).cuda().view(torch.uint16)))
        elif k % 3 == 1:
            imgs.append(cle.push(a.astype(np.float32)))
        else:
            imgs.append(cle.asarray(torch.from_numpy(a.astype(np.uint8)).cuda()))
        sig.append(float(rng.choice([0.8, 1.5, 2.0, 3.3, 5.6])))  # 5.6: 45 taps, the per-image path
    cle.set_autotune(False)
    base = [cle.pull(o) for o in cle.gaussian_blur_batch(imgs, sig)]
    cle.set_autotune(True)
    tuning = [cle.pull(o) for o in cle.gaussian_blur_batch(imgs, sig, autotune=True)]
    tuned = [cle.pull(o) for o in cle.gaussian_blur_batch(imgs, sig)]
    for k, (im, s) in enumerate(zip(imgs, sig)):
        single = cle.pull(cle.gaussian_blur(im, None, s, s, s))
        assert np.array_equal(base[k], single), k
        assert np.array_equal(tuning[k], single) and np.array_equal(tuned[k], single), k
        if k < 3:
            src = cle.pull(im).astype(np.float32)
            want = O.gaussian_blur(src, s, s, s)
            assert float(np.abs(base[k] - want).max()) <= 1e-4 * max(1.0, float(np.abs(want).max())), k


def test_batch_autotuned_labels_identical(cle):
    rng = np.random.default_rng(71)
    imgs, ss, so = _mixed_batch(rng, cle)
    cle.set_autotune(False)
    want = [cle.pull(o) for o in cle.voronoi_otsu_labeling_batch(imgs, ss, so)]
    cle.set_autotune(True)
    got1 = [cle.pull(o) for o in cle.autotune_voronoi_otsu_labeling_batch(imgs, ss, so)]
    got2 = [cle.pull(o) for o in cle.voronoi_otsu_labeling_batch(imgs, ss, so)]
    for k in range(len(imgs)):
        assert np.array_equal(got1[k], want[k]) and np.array_equal(got2[k], want[k]), k
