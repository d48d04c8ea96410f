"""Device images: the Array type and push / pull / create*.

An Array owns a contiguous PyTorch-ROCm tensor on the current HIP device (PyTorch
provides the allocation and the stream; every computation runs in libcle_hip.so).
Arrays exchange data zero-copy through DLPack (__dlpack__ / __dlpack_device__).

Reference: pyclesperanto_prototype/_tier0/_push.py:4-41, _pull.py:4-28,
_create.py:5-160 and the OCLArray type (_tier0/_pycl.py:63-182).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the runtime
    torch = None

_NP_TO_TORCH = None


def _torch_dtype(dt: np.dtype):
    global _NP_TO_TORCH
    if _NP_TO_TORCH is None:
        _NP_TO_TORCH = {
            np.dtype(np.uint8): torch.uint8, np.dtype(np.int8): torch.int8,
            np.dtype(np.uint16): torch.uint16, np.dtype(np.int16): torch.int16,
            np.dtype(np.uint32): torch.uint32, np.dtype(np.int32): torch.int32,
            np.dtype(np.float32): torch.float32,
            np.dtype(np.int64): torch.int64, np.dtype(np.float64): torch.float64,
        }
    try:
        return _NP_TO_TORCH[np.dtype(dt)]
    except KeyError:
        raise KeyError(f"dtype {np.dtype(dt)} not supported") from None


def _np_dtype(tdt) -> np.dtype:
    for k, v in (_NP_TO_TORCH or {}).items():
        if v == tdt:
            return k
    _torch_dtype(np.float32)
    for k, v in _NP_TO_TORCH.items():
        if v == tdt:
            return k
    raise KeyError(f"torch dtype {tdt} not supported")


def device():
    if torch is None or not torch.cuda.is_available():
        raise RuntimeError("no HIP device available: pyclesperanto_prototype_amd needs an AMD GPU (gfx950)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle() -> ctypes.c_void_p:
    """The hipStream_t of PyTorch's current stream."""
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Array:
    """A contiguous device image (z, y, x) / (y, x) / (n,)."""

    __slots__ = ("_t", "__weakref__")

    def __init__(self, tensor):
        # kernels dereference the data pointer on the GPU: host memory must never get here
        if tensor.device.type != "cuda":
            raise ValueError(f"cle.Array needs a device tensor, got one on {tensor.device}; use cle.push / cle.asarray")
        if not tensor.is_contiguous():
            tensor = tensor.contiguous()
        self._t = tensor

    # ---------------------------------------------------------------- metadata
    @property
    def shape(self):
        return tuple(self._t.shape)

    @property
    def dtype(self) -> np.dtype:
        return _np_dtype(self._t.dtype)

    @property
    def ndim(self) -> int:
        return self._t.dim()

    @property
    def size(self) -> int:
        return self._t.numel()

    @property
    def nbytes(self) -> int:
        return self._t.numel() * self._t.element_size()

    @property
    def tensor(self):
        """The underlying torch tensor (shares memory)."""
        return self._t

    def data_ptr(self) -> int:
        return self._t.data_ptr()

    def buf(self) -> _lib.cle_buf:
        """The C-ABI descriptor of this image."""
        code = _lib.NP_TO_CODE.get(self.dtype)
        if code is None:
            raise TypeError(f"dtype {self.dtype} is not supported by the HIP kernels")
        b = _lib.cle_buf()
        b.data = self._t.data_ptr()
        b.dtype = code
        shp = self.shape if self.ndim else (1,)
        if len(shp) > 3:
            raise ValueError(f"images have at most 3 dimensions, got {len(shp)}")
        b.ndim = len(shp)
        st = self._t.stride() if self.ndim else (1,)
        for i, s in enumerate(shp):
            b.shape[i] = s
            b.strides[i] = st[i]
        return b

    # ---------------------------------------------------------------- transfer
    def get(self) -> np.ndarray:
        t = self._t
        if t.dtype in (torch.uint16, torch.uint32):
            # copy the raw bits through a same-width signed type
            sig = torch.int16 if t.dtype == torch.uint16 else torch.int32
            return t.view(sig).cpu().numpy().view(self.dtype)
        return t.cpu().numpy()

    def __array__(self, dtype=None, copy=None):
        a = self.get()
        return a.astype(dtype) if dtype is not None else a

    def __dlpack__(self, *args, **kwargs):
        return self._t.__dlpack__(*args, **kwargs)

    def __dlpack_device__(self):
        return self._t.__dlpack_device__()

    def astype(self, dtype) -> "Array":
        out = create(self.shape, dtype)
        _lib.call("cle_copy", self.buf(), out.buf(), stream_handle())
        return out

    def copy(self) -> "Array":
        return Array(self._t.clone())

    # ---------------------------------------------------------------- reductions
    def min(self, axis=None, out=None):
        if axis is not None:
            raise NotImplementedError("axis reductions are not part of this build")
        return _minmax(self)[0]

    def max(self, axis=None, out=None):
        if axis is not None:
            raise NotImplementedError("axis reductions are not part of this build")
        return _minmax(self)[1]

    def sum(self, axis=None, out=None):
        if axis is not None:
            raise NotImplementedError("axis reductions are not part of this build")
        res = torch.empty(1, dtype=torch.float64, device=self._t.device)
        _lib.call("cle_sum_of_all_pixels", self.buf(), ctypes.c_void_p(res.data_ptr()), stream_handle())
        return np.float32(res.item())

    def __repr__(self):
        return f"cle.Array({self.get()!r}, dtype={self.dtype})"

    def __len__(self):
        return self.shape[0]


def _minmax(a: Array):
    res = torch.empty(2, dtype=torch.float32, device=a.tensor.device)
    _lib.call("cle_minimum_maximum", a.buf(), ctypes.c_void_p(res.data_ptr()), stream_handle())
    v = res.cpu().numpy()
    return np.float32(v[0]), np.float32(v[1])


def _from_numpy(a: np.ndarray) -> Array:
    a = np.ascontiguousarray(a)
    dt = a.dtype
    if dt in (np.dtype(np.uint16), np.dtype(np.uint32)):
        sig = np.int16 if dt == np.dtype(np.uint16) else np.int32
        t = torch.from_numpy(a.view(sig)).to(device()).view(_torch_dtype(dt))
    else:
        t = torch.from_numpy(a).to(device())
    return Array(t)


def asarray(x) -> Array:
    """Device Array of `x` keeping its element type: an Array or device tensor is
    wrapped without a copy, other DLPack producers are imported (and uploaded if they
    live in host memory), numpy / array-likes are uploaded."""
    if isinstance(x, Array):
        return x
    if isinstance(x, (np.ndarray, list, tuple)) or np.isscalar(x):
        return _from_numpy(np.asarray(x))
    if torch is not None and isinstance(x, torch.Tensor):
        return Array(x if x.device.type == "cuda" else x.to(device()))
    if hasattr(x, "__dlpack__"):
        t = torch.from_dlpack(x)
        return Array(t if t.device.type == "cuda" else t.to(device()))
    return _from_numpy(np.asarray(x))


def push(any_array) -> Array:
    """Copy an image to device memory as float32 (an Array passes through unchanged).

    Reference: pyclesperanto_prototype/_tier0/_push.py:28-41.  A device tensor is
    converted on the device instead of the reference's round trip through the host.
    """
    if isinstance(any_array, Array):
        return any_array
    if torch is not None and isinstance(any_array, torch.Tensor):
        t = any_array.to(device=device(), dtype=torch.float32)
        return Array(t.contiguous())
    if isinstance(any_array, (list, tuple)):
        any_array = np.asarray(any_array)
    if hasattr(any_array, "shape") and hasattr(any_array, "dtype") and hasattr(any_array, "get"):
        any_array = np.asarray(any_array.get())
    return _from_numpy(np.asarray(any_array, dtype=np.float32))


def push_zyx(any_array) -> Array:
    return push(any_array)


def pull(gpu_array) -> np.ndarray:
    """Reference: pyclesperanto_prototype/_tier0/_pull.py:4-28."""
    if isinstance(gpu_array, Array):
        return gpu_array.get()
    return asarray(gpu_array).get()


def pull_zyx(gpu_array) -> np.ndarray:
    return pull(gpu_array)


def create(dimensions, dtype=np.float32) -> Array:
    """Uninitialised device image.  Reference: pyclesperanto_prototype/_tier0/_create.py:5-23."""
    if isinstance(dimensions, Array):
        dimensions = dimensions.shape
    dims = tuple(int(d) for d in dimensions)
    return Array(torch.empty(dims, dtype=_torch_dtype(np.dtype(dtype)), device=device()))


def create_zyx(dimensions) -> Array:
    return create(dimensions[::-1])


def _dims_of(x, reverse_ndarray: bool):
    if isinstance(x, Array):
        return x.shape
    if isinstance(x, np.ndarray):
        return x.shape[::-1] if reverse_ndarray else x.shape
    if hasattr(x, "shape"):
        return tuple(x.shape)
    return x


def create_like(*args) -> Array:
    """float32 image of the same shape (_create.py:25-32)."""
    return create(_dims_of(args[0], False))


def create_binary_like(*args) -> Array:
    """uint8 image (_create.py:34-42; a raw numpy argument has its shape reversed there too)."""
    return create(_dims_of(args[0], True), np.uint8)


def create_labels_like(*args) -> Array:
    """uint32 image (_create.py:44-52; a raw numpy argument has its shape reversed there too)."""
    return create(_dims_of(args[0], True), np.uint32)


def create_none(*args):
    return None
