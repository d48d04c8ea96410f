"""cle_voronoi_otsu_labeling_noalloc against cle_voronoi_otsu_labeling through the C ABI:
identical labels on two calls in a row (the second on a smaller image, so it runs on
what the first left in the shared scratch array), no device memory taken from the
driver by those calls, a clear error and an untouched destination for an image beyond
the fixed scratch, and the oracle on a small image."""
import ctypes

import numpy as np
import pytest

from oracle import cle_oracle as O

pytestmark = pytest.mark.gpu

CLE_E_VALUE = -3


def _noalloc(cle, src, dst, spot, outline):
    from pyclesperanto_prototype_amd import _lib
    return _lib.call_rc("cle_voronoi_otsu_labeling_noalloc", src.buf(), dst.buf(), ctypes.c_float(spot),
                        ctypes.c_float(outline), cle._ops._S())


def _plain(cle, src, dst, spot, outline):
    from pyclesperanto_prototype_amd import _lib
    _lib.call("cle_voronoi_otsu_labeling", src.buf(), dst.buf(), ctypes.c_float(spot), ctypes.c_float(outline),
              cle._ops._S())


def test_noalloc_matches_the_allocating_entry_point(cle):
    import torch
    from pyclesperanto_prototype_amd import _lib, synthetic
    big_t, _ = synthetic.nuclei_3d((64, 256, 256), n_nuclei=300, seed=11)
    big = cle.asarray(big_t)
    small = cle.push(synthetic.blobs_2d((512, 512), seed=5))
    cases = [(big, 2.0, 1.5), (small, 3.0, 1.0)]
    outs = [cle.create(a.shape, np.uint32) for a, _, _ in cases]
    warm = cle.create(small.shape, np.uint32)
    torch.cuda.synchronize()
    assert _noalloc(cle, small, warm, 3.0, 1.0) == 0, _lib.last_error()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for (src, spot, outline), out in zip(cases, outs):
        print(f"noalloc call {tuple(src.shape)} spot={spot} outline={outline}", flush=True)
        assert _noalloc(cle, src, out, spot, outline) == 0, _lib.last_error()
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    for (src, spot, outline), out in zip(cases, outs):
        ref = cle.create(src.shape, np.uint32)
        _plain(cle, src, ref, spot, outline)
        got, want = cle.pull(out), cle.pull(ref)
        assert want.max() > 10
        assert np.array_equal(got, want), (tuple(src.shape), int(np.sum(got != want)))


def test_noalloc_rejects_an_image_beyond_its_scratch(cle):
    import torch
    from pyclesperanto_prototype_amd import _lib
    src = cle.create((2700, 2700), np.uint8)
    dst = cle.create((2700, 2700), np.uint32)
    cle.set(dst, 7)
    torch.cuda.synchronize()
    rc = _noalloc(cle, src, dst, 2.0, 2.0)
    assert rc == CLE_E_VALUE
    assert "noalloc" in _lib.last_error() and "scratch" in _lib.last_error()
    assert np.all(cle.pull(dst) == 7)


def test_noalloc_near_its_limit_and_the_oracle(cle):
    from pyclesperanto_prototype_amd import _lib, synthetic
    # 2100 x 2100 (4.41 M voxels) still fits the fixed scratch; labels as the allocating call
    img = synthetic.blobs_2d((2100, 2100), n_spots=700, seed=8)
    src = cle.push(img)
    out, ref = cle.create(img.shape, np.uint32), cle.create(img.shape, np.uint32)
    print("noalloc call (2100, 2100)", flush=True)
    assert _noalloc(cle, src, out, 2.0, 2.0) == 0, _lib.last_error()
    _plain(cle, src, ref, 2.0, 2.0)
    assert np.array_equal(cle.pull(out), cle.pull(ref))
    # a small image against the oracle
    small = synthetic.blobs_2d((200, 180), n_spots=20, seed=9)
    s = cle.push(small)
    o = cle.create(small.shape, np.uint32)
    assert _noalloc(cle, s, o, 2.0, 2.0) == 0, _lib.last_error()
    want = O.voronoi_otsu_labeling(small, 2.0, 2.0)
    assert np.array_equal(cle.pull(o), want)
