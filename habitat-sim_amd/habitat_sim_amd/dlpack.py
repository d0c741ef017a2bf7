"""Zero-copy export of the renderer's device planes through DLPack.

The reference hands sensor observations to PyTorch on the GPU
(``gpu2gpu_transfer``, src_python/habitat_sim/sensors/sensor_wrapper.py:
97-124) by copying the GL framebuffer into a torch CUDA tensor.  Here the
planes already live in device memory in observation layout (one image per
environment, consecutive), so they are exported as DLPack tensors of device
type ROCm and ``torch.from_dlpack`` wraps them without a copy.  A view is
valid until the renderer's next ``draw()`` writes the planes again (clone
it to keep a frame) and must not outlive the renderer.

With torch, initialise torch's GPU before libgbr loads: torch brings its own
HIP runtime and libgbr binds to the one already in the process.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import numpy as np

KDL_ROCM = 10  # DLDeviceType kDLROCM
_KDL_INT, _KDL_UINT, _KDL_FLOAT = 0, 1, 2
_DTYPES = {np.dtype(np.uint8): (_KDL_UINT, 8), np.dtype(np.int32): (_KDL_INT, 32),
           np.dtype(np.uint32): (_KDL_UINT, 32), np.dtype(np.float32): (_KDL_FLOAT, 32)}


class _DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int32), ("device_id", ctypes.c_int32)]


class _DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class _DLTensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("device", _DLDevice), ("ndim", ctypes.c_int32),
                ("dtype", _DLDataType), ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


class _DLManagedTensor(ctypes.Structure):
    pass


_DLManagedTensor._fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", ctypes.c_void_p),
                             ("deleter", ctypes.c_void_p)]  # void (*)(DLManagedTensor*)

# The exported DLManagedTensor and its shape live in one malloc'd block whose
# deleter is libc's free: the consumer may drop the tensor during interpreter
# shutdown, when a Python callback could no longer run.  The owner is not
# kept alive by the tensor (like the reference's raw device pointers): a view
# must not outlive its renderer.
_libc = ctypes.CDLL(None)
_libc.malloc.restype = ctypes.c_void_p
_libc.malloc.argtypes = [ctypes.c_size_t]
_FREE = ctypes.cast(_libc.free, ctypes.c_void_p).value

_pyapi = ctypes.pythonapi
_pyapi.PyCapsule_New.restype = ctypes.py_object
_pyapi.PyCapsule_New.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]


class DeviceArray:
    """Device memory of `shape` and `dtype` at `ptr` on HIP device
    `device_id`, exported through DLPack; `owner` (the renderer) stays alive
    as long as this object, not as long as the tensors made from it."""

    def __init__(self, ptr: int, shape: Tuple[int, ...], dtype, device_id: int, owner,
                 sync=None):
        self.ptr = int(ptr)
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        if self.dtype not in _DTYPES:
            raise TypeError(f"no DLPack type for {self.dtype}")
        self.device_id = int(device_id)
        self.owner = owner
        self._sync = sync

    def __dlpack_device__(self):
        return (KDL_ROCM, self.device_id)

    def __dlpack__(self, stream: Optional[int] = None, **_):
        # the planes are complete once the producer's stream is drained
        if self._sync is not None:
            self._sync()
        nd = len(self.shape)
        size = ctypes.sizeof(_DLManagedTensor)
        block = _libc.malloc(size + 8 * max(nd, 1))
        if not block:
            raise MemoryError("DLPack export")
        m = _DLManagedTensor.from_address(block)
        shape = (ctypes.c_int64 * max(nd, 1)).from_address(block + size)
        for k, v in enumerate(self.shape):
            shape[k] = v
        code, bits = _DTYPES[self.dtype]
        t = m.dl_tensor
        t.data = self.ptr
        t.device = _DLDevice(KDL_ROCM, self.device_id)
        t.ndim = nd
        t.dtype = _DLDataType(code, bits, 1)
        t.shape = ctypes.cast(block + size, ctypes.POINTER(ctypes.c_int64))
        t.strides = None  # compact row-major
        t.byte_offset = 0
        m.manager_ctx = None
        m.deleter = _FREE
        # an unconsumed capsule leaks its block (a few dozen bytes)
        return _pyapi.PyCapsule_New(block, b"dltensor", None)
