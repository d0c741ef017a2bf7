"""Redwood depth noise on the GPU: habitat-sim's RedwoodDepthNoiseModel
(src_python/habitat_sim/sensors/noise_models/redwood_depth_noise_model.py:107-148)
over the HIP kernel of libgbr.so (the reference's RedwoodNoiseModelGPUImpl,
src/esp/sensor/RedwoodNoiseModel.{h,cu}).

``simulate(depth)`` takes a host array (H, W) and returns a host array, or a
torch tensor on the GPU, (H, W) or a batch (N, H, W), and returns a new
tensor on the same device, like the reference.  ``simulate_device`` works on
raw device pointers, e.g. the batch renderer's linear-depth plane (one
image per environment, consecutive).

With torch tensors, initialise torch's GPU before this library loads
(``torch.cuda.init()``): torch brings its own HIP runtime and libgbr then
binds to the one already in the process; the other order leaves torch
without a device.

The undistortion table is the reference's ``redwood-depth-dist-model.npy``
(80 x 80 x 5 floats), which the caller passes (``load_model(path)`` reads
it).  Deviates come from a Philox4x32-10 counter generator keyed by
``seed``: call k of a model reproduces from (seed, k).
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _native as N


def load_model(path: str) -> np.ndarray:
    """The (80, 80, 5) float32 undistortion table from a .npy file."""
    m = np.load(path, allow_pickle=False)
    return np.ascontiguousarray(m, np.float32).reshape(80, 80, 5)


class RedwoodDepthNoiseModel:
    def __init__(self, model, noise_multiplier: float = 1.0, gpu_device_id: int = -1,
                 seed: int = 0):
        m = np.ascontiguousarray(np.asarray(model, np.float32).reshape(-1))
        if m.size != N.REDWOOD_MODEL_FLOATS:
            raise ValueError(f"the undistortion table has {m.size} values, "
                             f"expected {N.REDWOOD_MODEL_FLOATS} (80 x 80 x 5)")
        self.noise_multiplier = float(noise_multiplier)
        h = ctypes.c_void_p()
        N.call("gbr_redwood_create", m.ctypes.data, int(gpu_device_id), self.noise_multiplier,
               int(seed) & 0xFFFFFFFFFFFFFFFF, ctypes.byref(h))
        self._h = h

    def close(self) -> None:
        if self._h:
            N.call("gbr_redwood_destroy", self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def calls(self) -> int:
        """Generator calls so far (the index of the next one)."""
        out = ctypes.c_uint64()
        N.call("gbr_redwood_calls", self._h, ctypes.byref(out))
        return out.value

    def simulate(self, gt_depth):
        if isinstance(gt_depth, np.ndarray):
            d = np.ascontiguousarray(gt_depth, np.float32)
            if d.ndim != 2:
                raise ValueError("host depth must be (H, W)")
            out = np.empty_like(d)
            N.call("gbr_redwood_simulate_host", self._h, d.ctypes.data, d.shape[0], d.shape[1],
                   out.ctypes.data)
            return out
        import torch
        if not (isinstance(gt_depth, torch.Tensor) and gt_depth.is_cuda):
            raise TypeError("depth must be a numpy array or a GPU torch tensor")
        d = gt_depth.contiguous().to(torch.float32)
        if d.dim() not in (2, 3):
            raise ValueError("device depth must be (H, W) or (N, H, W)")
        out = torch.empty_like(d)
        n = 1 if d.dim() == 2 else d.shape[0]
        stream = torch.cuda.current_stream(d.device).cuda_stream
        self.simulate_device(d.data_ptr(), n, d.shape[-2], d.shape[-1], out.data_ptr(), stream)
        return out

    apply = simulate
    __call__ = simulate

    def simulate_device(self, depth_ptr: int, images: int, rows: int, cols: int, out_ptr: int,
                        stream: Optional[int] = None) -> None:
        """simulateFromGPU over `images` consecutive rows x cols images."""
        N.call("gbr_redwood_simulate", self._h, depth_ptr, images, rows, cols, out_ptr,
               stream or None)

    def simulate_device_normals(self, depth_ptr: int, images: int, rows: int, cols: int,
                                normals_ptr: int, out_ptr: int, stream: Optional[int] = None) -> None:
        """The same with caller-given N(0, 1) deviates, 3 per pixel."""
        N.call("gbr_redwood_simulate_normals", self._h, depth_ptr, images, rows, cols,
               normals_ptr, out_ptr, stream or None)

    def random_bits_device(self, call: int, first: int, n: int, out_ptr: int,
                           stream: Optional[int] = None) -> None:
        """The generator's raw Philox4x32-10 blocks (4 words each)."""
        N.call("gbr_redwood_random_bits", self._h, call, first, n, out_ptr, stream or None)
