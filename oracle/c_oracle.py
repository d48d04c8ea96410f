"""ctypes wrapper of oracle/lib/liboracle.so (the C restatement) -- test
infrastructure and CPU baseline only."""
import ctypes
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "lib" / "liboracle.so"
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_I = ctypes.c_int64


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            from oracle import build_c
            build_c.build(verbose=False)
        l = ctypes.CDLL(str(LIB))
        l.ora_gaussian_blur.argtypes = [_f32p, _f32p, _I, _I, _I, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float]
        l.ora_detect_maxima.argtypes = [_f32p, _u8p, _I, _I, _I, ctypes.c_int]
        l.ora_threshold_otsu.argtypes = [_f32p, _I, _I, _I, _u8p]
        l.ora_threshold_otsu.restype = ctypes.c_double
        l.ora_ccl.argtypes = [_u8p, _I, _I, _I, ctypes.c_int, ctypes.c_int, _f32p]
        l.ora_ccl.restype = ctypes.c_int
        l.ora_masked_voronoi.argtypes = [_u8p, _u8p, _I, _I, _I, ctypes.c_int, _u32p]
        l.ora_masked_voronoi.restype = ctypes.c_int
        l.ora_voronoi_otsu_labeling.argtypes = [_f32p, _u32p, _I, _I, _I, ctypes.c_int, ctypes.c_float, ctypes.c_float]
        l.ora_statistics.argtypes = [_f32p, _u32p, _I, _I, _I, _I, _f32p]
        l.ora_blur_otsu_ccl_stats.argtypes = [_f32p, _I, _I, _I, ctypes.c_int, ctypes.c_float, _i32p, _f32p, _I]
        l.ora_blur_otsu_ccl_stats.restype = ctypes.c_int64
        _lib = l
    return _lib


def set_threads(n):
    """OpenMP threads of the C restatement's parallel loops."""
    lib()
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))


def _dims(a):
    return (1,) + a.shape if a.ndim == 2 else a.shape


def gaussian_blur(a, sx, sy, sz=0.0):
    a = np.ascontiguousarray(a, np.float32)
    out = np.empty_like(a)
    lib().ora_gaussian_blur(a, out, *_dims(a), a.ndim, sx, sy, sz)
    return out


def detect_maxima(a):
    a = np.ascontiguousarray(a, np.float32)
    out = np.empty(a.shape, np.uint8)
    lib().ora_detect_maxima(a, out, *_dims(a), a.ndim)
    return out


def threshold_otsu(a):
    a = np.ascontiguousarray(a, np.float32)
    out = np.empty(a.shape, np.uint8)
    t = lib().ora_threshold_otsu(a, *_dims(a), out)
    return out, t


def ccl(binary, diamond=False):
    b = np.ascontiguousarray(binary != 0, np.uint8)
    out = np.empty(b.shape, np.float32)
    it = lib().ora_ccl(b, *_dims(b), b.ndim, int(diamond), out)
    return out.astype(np.uint32), it


def masked_voronoi(binary, mask):
    b = np.ascontiguousarray(binary != 0, np.uint8)
    m = np.ascontiguousarray(mask != 0, np.uint8)
    out = np.empty(b.shape, np.uint32)
    it = lib().ora_masked_voronoi(b, m, *_dims(b), b.ndim, out)
    return out, it


def voronoi_otsu_labeling(a, ss, so):
    a = np.ascontiguousarray(a, np.float32)
    out = np.empty(a.shape, np.uint32)
    lib().ora_voronoi_otsu_labeling(a, out, *_dims(a), a.ndim, ss, so)
    return out


def statistics_table(img, labels):
    img = np.ascontiguousarray(img, np.float32)
    lab = np.ascontiguousarray(labels, np.uint32)
    L = int(lab.max()) + 1
    table = np.zeros((36, max(L - 1, 1)), np.float32)
    lib().ora_statistics(img, lab, *_dims(lab), L, table)
    return table[:, : L - 1]


def blur_otsu_ccl_stats(a, sigma, with_stats=True):
    a = np.ascontiguousarray(a, np.float32)
    labels = np.empty(a.shape, np.int32)
    cap = 36 * 200000 if with_stats else 0  # up to 200k labels
    table = np.zeros(max(cap, 1), np.float32)
    L = lib().ora_blur_otsu_ccl_stats(a, *_dims(a), a.ndim, sigma, labels, table if with_stats else table, cap)
    return labels, (table[: 36 * (L - 1)].reshape(36, L - 1) if with_stats else None)
