"""INTEGRATION.md's reference-side ctypes stub, executed as written (its two python
blocks) against the built library: every stub function gives what the package gives
for the same input, the numpy path (cle_alloc / cle_upload / cle_download / cle_free)
included, and a strided view is refused with ValueError (CLE_E_SHAPE)."""
import os
import re
import types

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stub(cle):
    from pyclesperanto_prototype_amd import _lib
    text = (ROOT / "INTEGRATION.md").read_text()
    blocks = re.findall(r"```python\n(.*?)```", text, re.S)
    stub = [b for b in blocks if b.startswith("# pyclesperanto_prototype/_tier0/_hip_hotpath.py")]
    numpy_part = [b for b in blocks if b.startswith("# continues the stub above")]
    assert len(stub) == 1 and len(numpy_part) == 1, "INTEGRATION.md lost its stub"
    os.environ["CLE_HIP_LIB"] = str(_lib.LIB_PATH)
    mod = types.ModuleType("hip_hotpath_stub")
    exec(compile(stub[0] + "\n" + numpy_part[0], "INTEGRATION.md", "exec"), mod.__dict__)
    return mod


def _nuclei(shape, n, seed):
    from pyclesperanto_prototype_amd import synthetic
    t, _ = synthetic.nuclei_3d(shape, n, seed=seed)
    return t


def test_stub_voronoi_otsu_and_batch(cle, stub):
    import torch
    src = _nuclei((40, 160, 192), 60, 3)
    dst = torch.zeros(src.shape, dtype=torch.int32, device=src.device)
    stub.voronoi_otsu_labeling(src, dst, 2.0, 1.5)
    want = cle.pull(cle.voronoi_otsu_labeling(cle.asarray(src), None, 2.0, 1.5))
    assert np.array_equal(dst.cpu().numpy(), want)
    assert want.max() > 10
    srcs = [src, _nuclei((24, 96, 128), 20, 4)]
    dsts = [torch.zeros(s.shape, dtype=torch.int32, device=s.device) for s in srcs]
    stub.voronoi_otsu_labeling_many(srcs, dsts, [2.0, 1.0], [1.5, 1.0])
    for s, d, a, b in zip(srcs, dsts, [2.0, 1.0], [1.5, 1.0]):
        assert np.array_equal(d.cpu().numpy(), cle.pull(cle.voronoi_otsu_labeling(cle.asarray(s), None, a, b)))


def test_stub_ccl_statistics_extend(cle, stub):
    import torch
    src = _nuclei((32, 128, 160), 40, 5)
    binary = (src.to(torch.int32) > 600).to(torch.uint8)
    lab = torch.zeros(src.shape, dtype=torch.int32, device=src.device)
    stub.connected_components_labeling_box(binary, lab)
    want = cle.pull(cle.connected_components_labeling_box(cle.asarray(binary)))
    assert np.array_equal(lab.cpu().numpy(), want)
    n = int(want.max()) + 1
    table = stub.statistics_table(src.to(torch.float32), lab, n)
    ref, nref = cle.statistics_table(cle.asarray(src.to(torch.float32)), cle.asarray(lab), n)
    assert nref == n - 1 and table.shape[1] == n - 1
    assert np.array_equal(table.cpu().numpy(), ref.tensor.cpu().numpy()[:, :n - 1])
    ext = torch.zeros(src.shape, dtype=torch.int32, device=src.device)
    stub.extend_labeling_via_voronoi(lab, ext)
    assert np.array_equal(ext.cpu().numpy(), cle.pull(cle.extend_labeling_via_voronoi(cle.asarray(lab))))


def test_stub_numpy_path(cle, stub):
    from pyclesperanto_prototype_amd import synthetic
    img = synthetic.blobs_2d((300, 260), seed=6)
    got = stub.voronoi_otsu_labeling_numpy(img, 3.0, 1.0)
    assert got.dtype == np.uint32
    assert np.array_equal(got, cle.pull(cle.voronoi_otsu_labeling(cle.push(img), None, 3.0, 1.0)))


def test_stub_refuses_strided_views(cle, stub):
    import torch
    src = _nuclei((16, 64, 64), 5, 7)
    view = src[:, :, ::2]
    dst = torch.zeros(view.shape, dtype=torch.int32, device=src.device)
    with pytest.raises(ValueError, match="contiguous"):
        stub.voronoi_otsu_labeling(view, dst)
