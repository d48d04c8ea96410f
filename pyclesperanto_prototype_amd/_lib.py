"""ctypes binding of libcle_hip.so (declared in include/cle_abi.h).

The library is loaded lazily on first use; a missing library raises ImportError
naming the build command.  There is no fallback path: every op runs the HIP kernels.
"""
from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libcle_hip.so"

# element type codes (cle_abi.h)
CLE_U8, CLE_I8, CLE_U16, CLE_I16, CLE_U32, CLE_I32, CLE_F32 = 1, 2, 3, 4, 5, 6, 7
NP_TO_CODE = {
    np.dtype(np.uint8): CLE_U8, np.dtype(np.int8): CLE_I8,
    np.dtype(np.uint16): CLE_U16, np.dtype(np.int16): CLE_I16,
    np.dtype(np.uint32): CLE_U32, np.dtype(np.int32): CLE_I32,
    np.dtype(np.float32): CLE_F32,
}

CLE_E_TYPE, CLE_E_SHAPE, CLE_E_VALUE, CLE_E_OOM, CLE_E_HIP = -1, -2, -3, -4, -5
CLE_CONN_BOX, CLE_CONN_DIAMOND = 0, 1
CLE_FP32_KEY_COMPAT = 1
CLE_STAT_COLUMNS = 36


class cle_buf(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("dtype", ctypes.c_int32), ("ndim", ctypes.c_int32),
                ("shape", ctypes.c_int64 * 3), ("strides", ctypes.c_int64 * 3)]


_P = ctypes.POINTER(cle_buf)
_V = ctypes.c_void_p
_F = ctypes.c_float
_I = ctypes.c_int

# name -> argtypes (restype int unless listed in _STR)
SIGNATURES = {
    "cle_version": [],
    "cle_last_error": [],
    "cle_release_pool": [],
    "cle_release_stream": [_V],
    "cle_reserve_stream": [_V, ctypes.c_size_t],
    "cle_alloc": [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)],
    "cle_free": [_V],
    "cle_upload": [_V, _V, ctypes.c_size_t, _V],
    "cle_download": [_V, _V, ctypes.c_size_t, _V],
    "cle_set": [_P, _F, _V],
    "cle_copy": [_P, _P, _V],
    "cle_gaussian_blur": [_P, _P, _F, _F, _F, _V],
    "cle_mean_box": [_P, _P, _I, _I, _I, _V],
    "cle_detect_maxima_box": [_P, _P, _I, _I, _I, _V],
    "cle_minimum_maximum": [_P, _V, _V],
    "cle_sum_of_all_pixels": [_P, _V, _V],
    "cle_histogram": [_P, _P, _I, _F, _F, _V, _V],
    "cle_threshold_otsu": [_P, _P, _V, _V],
    "cle_greater_constant": [_P, _P, _F, _V],
    "cle_binary_and": [_P, _P, _P, _V],
    "cle_mask": [_P, _P, _P, _V],
    "cle_add_image_and_scalar": [_P, _P, _F, _V],
    "cle_add_images_weighted": [_P, _P, _P, _F, _F, _V],
    "cle_set_nonzero_pixels_to_pixel_index": [_P, _P, _I, _V],
    "cle_nonzero_minimum_box": [_P, _P, _P, _V],
    "cle_nonzero_minimum_diamond": [_P, _P, _P, _V],
    "cle_onlyzero_overwrite_maximum_box": [_P, _P, _P, _V],
    "cle_onlyzero_overwrite_maximum_diamond": [_P, _P, _P, _V],
    "cle_connected_components_labeling": [_P, _P, _I, _I, _V, _V],
    "cle_relabel_sequential": [_P, _P, _V],
    "cle_label_spots": [_P, _P, _V],
    "cle_masked_voronoi_labeling": [_P, _P, _P, ctypes.POINTER(ctypes.c_int), _V],
    "cle_extend_labeling_via_voronoi": [_P, _P, ctypes.POINTER(ctypes.c_int), _V],
    "cle_set_autotune": [ctypes.c_int],
    "cle_set_growth_method": [ctypes.c_int],
    "cle_growth_counters": [ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_int],
    "cle_autotune_gaussian_blur": [_P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_float, _V],
    "cle_gaussian_blur_batch": [_P, _P, ctypes.POINTER(ctypes.c_float), ctypes.c_int, _V],
    "cle_autotune_gaussian_blur_batch": [_P, _P, ctypes.POINTER(ctypes.c_float), ctypes.c_int, _V],
    "cle_voronoi_otsu_labeling_batch_plan": [_P, _P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                             ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), _V],
    "cle_plan_run": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), _V],
    "cle_plan_destroy": [ctypes.c_void_p],
    "cle_voronoi_otsu_labeling": [_P, _P, _F, _F, _V],
    "cle_voronoi_otsu_labeling_noalloc": [_P, _P, _F, _F, _V],
    "cle_gauss_otsu_labeling": [_P, _P, _F, _V],
    "cle_voronoi_otsu_labeling_batch": [_P, _P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                        ctypes.c_int, ctypes.POINTER(ctypes.c_int), _V],
    "cle_autotune_voronoi_otsu_labeling_batch": [_P, _P, ctypes.POINTER(ctypes.c_float),
                                                 ctypes.POINTER(ctypes.c_float), ctypes.c_int, _V],
    "cle_statistics_of_labelled_pixels": [_P, _P, ctypes.c_int64, _P, ctypes.POINTER(ctypes.c_int64), _V],
}
_STR = {"cle_version", "cle_last_error"}

_lib = None
_lock = threading.Lock()


def load():
    """Load and bind the library (idempotent)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m pyclesperanto_prototype_amd._build`")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_char_p if name in _STR else ctypes.c_int
        _lib = lib
    return _lib


def last_error() -> str:
    return load().cle_last_error().decode()


def check(rc: int, what: str) -> None:
    """Map a C-ABI status to the reference's exception types."""
    if rc == 0:
        return
    msg = f"{what}: {last_error()}"
    if rc == CLE_E_TYPE:
        raise TypeError(msg)
    if rc in (CLE_E_SHAPE, CLE_E_VALUE):
        raise ValueError(msg)
    if rc == CLE_E_OOM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


def call_rc(name: str, *args) -> int:
    """Call an entry point and return its status without raising."""
    return getattr(load(), name)(*args)
