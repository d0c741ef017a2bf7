"""Host-side logic that needs no GPU: the Redwood model's argument checks
and table loading, and the DLPack export's tensor description (consumers
such as torch.from_dlpack read exactly these fields)."""
import ctypes
import os

import numpy as np
import pytest

from habitat_sim_amd import dlpack
from habitat_sim_amd.noise import RedwoodDepthNoiseModel, load_model

MODEL = os.path.join(os.path.dirname(__file__), "golden", "noise", "redwood-depth-dist-model.npy")


def test_noise_model_table_checks():
    m = load_model(MODEL)
    assert m.shape == (80, 80, 5) and m.dtype == np.float32
    with pytest.raises(ValueError):
        RedwoodDepthNoiseModel(np.zeros(10, np.float32))  # rejected before any device call


def _tensor_of(capsule):
    pyapi = ctypes.pythonapi
    pyapi.PyCapsule_GetPointer.restype = ctypes.c_void_p
    pyapi.PyCapsule_GetPointer.argtypes = [ctypes.py_object, ctypes.c_char_p]
    pyapi.PyCapsule_IsValid.restype = ctypes.c_int
    pyapi.PyCapsule_IsValid.argtypes = [ctypes.py_object, ctypes.c_char_p]
    assert pyapi.PyCapsule_IsValid(capsule, b"dltensor")
    return dlpack._DLManagedTensor.from_address(pyapi.PyCapsule_GetPointer(capsule, b"dltensor"))


@pytest.mark.parametrize("shape,dtype,code,bits", [((3, 4, 5, 4), np.uint8, 1, 8),
                                                   ((2, 8, 8), np.float32, 2, 32),
                                                   ((1, 2, 2), np.int32, 0, 32)])
def test_dlpack_description(shape, dtype, code, bits):
    synced = []
    a = dlpack.DeviceArray(0x7000_0000_1000, shape, dtype, 3, owner=None,
                           sync=lambda: synced.append(1))
    assert a.__dlpack_device__() == (dlpack.KDL_ROCM, 3)
    m = _tensor_of(a.__dlpack__())
    t = m.dl_tensor
    assert synced == [1]  # the producer's stream is drained first
    assert t.data == 0x7000_0000_1000 and t.byte_offset == 0
    assert (t.device.device_type, t.device.device_id) == (dlpack.KDL_ROCM, 3)
    assert t.ndim == len(shape) and [t.shape[k] for k in range(t.ndim)] == list(shape)
    assert not t.strides  # compact row-major
    assert (t.dtype.code, t.dtype.bits, t.dtype.lanes) == (code, bits, 1)
    assert m.deleter  # libc free of the block (the consumer calls it)
    with pytest.raises(TypeError):
        dlpack.DeviceArray(0x1000, (2,), np.float64, 0, None)
