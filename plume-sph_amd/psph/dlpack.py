"""Zero-copy torch tensors over device arrays the library owns.

``psph_device_view`` returns plain device addresses.  ``as_tensor`` wraps one
of them as a 1-D torch tensor on the HIP device through a DLPack capsule
(the exchange format torch.from_dlpack reads), without copying.  The
tensor does not own the memory: it is valid while the library keeps the
array at that address (see psph_device_view in include/psph.h).
"""
from __future__ import annotations

import ctypes as C

import torch

_K_DL_INT, _K_DL_UINT, _K_DL_FLOAT = 0, 1, 2
_K_DL_ROCM = 10


class _DLDevice(C.Structure):
    _fields_ = [("device_type", C.c_int32), ("device_id", C.c_int32)]


class _DLDataType(C.Structure):
    _fields_ = [("code", C.c_uint8), ("bits", C.c_uint8), ("lanes", C.c_uint16)]


class _DLTensor(C.Structure):
    _fields_ = [("data", C.c_void_p), ("device", _DLDevice), ("ndim", C.c_int32), ("dtype", _DLDataType),
                ("shape", C.POINTER(C.c_int64)), ("strides", C.POINTER(C.c_int64)),
                ("byte_offset", C.c_uint64)]


class _DLManagedTensor(C.Structure):
    pass


_DELETER = C.CFUNCTYPE(None, C.POINTER(_DLManagedTensor))
_DLManagedTensor._fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", C.c_void_p), ("deleter", _DELETER)]

_DTYPES = {
    torch.float64: (_K_DL_FLOAT, 64),
    torch.float32: (_K_DL_FLOAT, 32),
    torch.int32: (_K_DL_INT, 32),
    torch.uint32: (_K_DL_UINT, 32),
}

# the memory belongs to the library: nothing to free when torch lets go
_noop_deleter = _DELETER(lambda _p: None)

# The capsule structures must outlive every tensor made from them (torch calls
# their deleter when the tensor dies); they are a few dozen bytes each, so
# they are kept for the life of the process.
KEEP: list = []

_PyCapsule_New = C.pythonapi.PyCapsule_New
_PyCapsule_New.restype = C.py_object
_PyCapsule_New.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]


def as_tensor(ptr: int, n: int, dtype: torch.dtype, device_index: int, keep: list) -> torch.Tensor:
    """1-D tensor of n elements of dtype at device address ptr (no copy).

    The ctypes objects behind the capsule are appended to ``keep``, which the
    caller holds for as long as the tensor is used."""
    if dtype not in _DTYPES:
        raise TypeError(f"unsupported dtype {dtype}")
    code, bits = _DTYPES[dtype]
    shape = (C.c_int64 * 1)(n)
    m = _DLManagedTensor()
    m.dl_tensor.data = C.c_void_p(ptr)
    m.dl_tensor.device = _DLDevice(_K_DL_ROCM, device_index)
    m.dl_tensor.ndim = 1
    m.dl_tensor.dtype = _DLDataType(code, bits, 1)
    m.dl_tensor.shape = shape
    m.dl_tensor.strides = None
    m.dl_tensor.byte_offset = 0
    m.manager_ctx = None
    m.deleter = _noop_deleter
    keep.extend((m, shape))
    cap = _PyCapsule_New(C.addressof(m), b"dltensor", None)
    return torch.from_dlpack(cap)
