"""CPU-side checks of the boundary: the library loads and exports every symbol that
include/cle_abi.h declares, argument validation maps to the reference's exception
types, and the Python API exposes the reference's names.  No kernel is launched."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        names |= set(re.findall(r"^\s*(?:int|const char\*)\s+(cle_[a-z0-9_]+)\s*\(", h.read_text(), re.M))
    return sorted(names)


def test_header_declares_the_abi():
    names = declared_symbols()
    assert "cle_voronoi_otsu_labeling" in names and "cle_statistics_of_labelled_pixels" in names
    assert len(names) >= 29


def test_library_exports_every_declared_symbol():
    from pyclesperanto_prototype_amd import _lib
    lib = _lib.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_symbols()) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"
    assert b"gfx950" in lib.cle_version()


def _buf(dtype_code, shape, data=0x1000):
    from pyclesperanto_prototype_amd import _lib
    b = _lib.cle_buf()
    b.data = data
    b.dtype = dtype_code
    b.ndim = len(shape)
    for i, s in enumerate(shape):
        b.shape[i] = s
    return b


def test_validation_errors_map_to_reference_exceptions():
    # these calls fail in host-side validation, before any HIP call
    from pyclesperanto_prototype_amd import _lib
    with pytest.raises(TypeError):
        _lib.call("cle_gaussian_blur", _buf(99, (4, 4)), _buf(_lib.CLE_F32, (4, 4)), 1.0, 1.0, 0.0, None)
    with pytest.raises(ValueError):
        _lib.call("cle_gaussian_blur", _buf(_lib.CLE_F32, (4, 4)), _buf(_lib.CLE_F32, (4, 5)), 1.0, 1.0, 0.0, None)
    with pytest.raises(ValueError):
        _lib.call("cle_binary_and", _buf(_lib.CLE_U8, (2, 3)), _buf(_lib.CLE_U8, (3, 2)), _buf(_lib.CLE_U8, (2, 3)), None)
    with pytest.raises(ValueError):
        _lib.call("cle_connected_components_labeling", _buf(_lib.CLE_U8, (2, 3)), _buf(_lib.CLE_U32, (2, 3)), 7, 0,
                  None, None)
    with pytest.raises(TypeError):
        tab = _buf(_lib.CLE_U32, (36, 4))
        _lib.call("cle_statistics_of_labelled_pixels", _buf(_lib.CLE_F32, (3, 3)), _buf(_lib.CLE_U32, (3, 3)), 5,
                  tab, None, None)
    assert "differ in shape" in _lib.last_error() or "float32" in _lib.last_error()
    with pytest.raises(ValueError):  # table rows must be the 36 STATISTICS_ENTRY columns
        _lib.call("cle_statistics_of_labelled_pixels", _buf(_lib.CLE_F32, (3, 3)), _buf(_lib.CLE_U32, (3, 3)), 5,
                  _buf(_lib.CLE_F32, (35, 4)), None, None)
    with pytest.raises(ValueError):  # explicit label count larger than the table
        _lib.call("cle_statistics_of_labelled_pixels", _buf(_lib.CLE_F32, (3, 3)), _buf(_lib.CLE_U32, (3, 3)), 9,
                  _buf(_lib.CLE_F32, (36, 4)), None, None)
    with pytest.raises(ValueError):
        _lib.call("cle_statistics_of_labelled_pixels", _buf(_lib.CLE_F32, (3, 3)), _buf(_lib.CLE_U32, (3, 3)), -1,
                  _buf(_lib.CLE_F32, (36, 4)), ctypes.byref(ctypes.c_int64(0)), None)


def test_api_surface_matches_reference_names():
    import pyclesperanto_prototype_amd as cle
    for name in ("push", "pull", "create", "create_like", "create_binary_like", "create_labels_like",
                 "plugin_function", "gaussian_blur", "detect_maxima_box", "threshold_otsu", "histogram",
                 "binary_and", "mask", "masked_voronoi_labeling", "voronoi_otsu_labeling",
                 "connected_components_labeling_box", "relabel_sequential", "label_spots",
                 "statistics_of_labelled_pixels", "statistics_of_labeled_pixels", "label",
                 "close_index_gaps_in_label_map", "set_nonzero_pixels_to_pixelindex", "logical_and",
                 "binary_intersection", "minimum_of_all_pixels", "maximum_of_all_pixels",
                 "extend_labeling_via_voronoi", "voronoi_labeling", "gauss_otsu_labeling"):
        assert callable(getattr(cle, name)), name
    assert cle.statistics_of_labeled_pixels is cle.statistics_of_labelled_pixels


def test_plugin_signature_binding_errors():
    import pyclesperanto_prototype_amd as cle
    with pytest.raises(TypeError):
        cle.gaussian_blur(np.zeros((2, 2)), None, 1, 1, 1, "extra", "args")
    with pytest.raises(TypeError):
        cle.voronoi_otsu_labeling(np.zeros((2, 2)), spot_size=3)
    with pytest.raises(NotImplementedError):
        cle.execute(None, "x.cl", "k", (1,), {})


def test_array_refuses_host_memory():
    """Kernels dereference Array data on the GPU; a host tensor must be rejected."""
    import torch
    import pyclesperanto_prototype_amd as cle
    with pytest.raises(ValueError):
        cle.Array(torch.zeros(4))
    with pytest.raises(ValueError):
        cle.Array(torch.from_dlpack(np.zeros(3, np.float32)))


def test_strided_descriptors_are_refused_on_the_host():
    # a non-C-contiguous stride fails in host validation (CLE_E_SHAPE), before any HIP call;
    # zero strides and the C-contiguous ones are accepted there
    from pyclesperanto_prototype_amd import _lib
    src, dst = _buf(_lib.CLE_F32, (4, 6)), _buf(_lib.CLE_F32, (4, 6))
    src.strides[0], src.strides[1] = 12, 2
    with pytest.raises(ValueError, match="contiguous"):
        _lib.call("cle_gaussian_blur", src, dst, 1.0, 1.0, 0.0, None)
    src.strides[0], src.strides[1] = 6, 1
    dst.shape[1] = 5  # fails later, at the shape comparison: the strides passed
    with pytest.raises(ValueError, match="differ in shape"):
        _lib.call("cle_gaussian_blur", src, dst, 1.0, 1.0, 0.0, None)


def test_alloc_of_zero_bytes_and_null_output():
    import ctypes
    from pyclesperanto_prototype_amd import _lib
    p = ctypes.c_void_p(123)
    _lib.call("cle_alloc", 0, ctypes.byref(p))
    assert p.value is None
    _lib.call("cle_free", None)
    with pytest.raises(ValueError):
        _lib.call("cle_alloc", 16, None)
