"""gbr_redwood_simulate_host_noalloc (the entry point that allocates nothing)
against gbr_redwood_simulate_host, through the C ABI: two models with the
same seed draw the same deviates call for call, so the outputs must be equal
bit for bit, on two calls in a row (the second with a smaller image, which
leaves the first call's pixels in the shared staging array beyond it).
Sizes above GBR_REDWOOD_NOALLOC_MAX_PIXELS are refused before any GPU work."""
import ctypes
import os

import numpy as np
import pytest

from habitat_sim_amd import _native as N
from habitat_sim_amd.noise import load_model

pytestmark = pytest.mark.gpu

MODEL = os.path.join(os.path.dirname(__file__), "golden", "noise", "redwood-depth-dist-model.npy")
MAX_PIXELS = 1 << 20


def _create(seed):
    h = ctypes.c_void_p()
    model = np.ascontiguousarray(load_model(MODEL), np.float32)
    N.call("gbr_redwood_create", model.ctypes.data, -1, 1.0, seed, ctypes.byref(h))
    return h


def _run(fn, h, depth):
    out = np.full(depth.shape, np.nan, np.float32)
    N.call(fn, h, depth.ctypes.data, depth.shape[0], depth.shape[1], out.ctypes.data)
    return out


def _calls(h):
    c = ctypes.c_uint64()
    N.call("gbr_redwood_calls", h, ctypes.byref(c))
    return c.value


def test_noalloc_matches_allocating_entry_point():
    rng = np.random.default_rng(5)
    shapes = [(480, 640), (37, 53)]  # second call smaller than the first
    a, b = _create(0xC0FFEE), _create(0xC0FFEE)
    try:
        for k, shape in enumerate(shapes):
            depth = rng.uniform(0.0, 12.0, shape).astype(np.float32)
            depth.reshape(-1)[::101] = np.nan
            ref = _run("gbr_redwood_simulate_host", a, depth)
            got = _run("gbr_redwood_simulate_host_noalloc", b, depth)
            np.testing.assert_array_equal(got, ref, err_msg=f"call {k}, shape {shape}")
            assert np.isfinite(got).all()
            assert _calls(a) == _calls(b) == k + 1
        # the largest image it takes, then one pixel row too many
        depth = rng.uniform(0.0, 9.0, (1024, 1024)).astype(np.float32)
        np.testing.assert_array_equal(_run("gbr_redwood_simulate_host_noalloc", b, depth),
                                      _run("gbr_redwood_simulate_host", a, depth))
        too_big = np.zeros((1025, 1024), np.float32)
        with pytest.raises(N.GbrError) as e:
            _run("gbr_redwood_simulate_host_noalloc", b, too_big)
        assert "GBR_REDWOOD_NOALLOC_MAX_PIXELS" in str(e.value)
        assert _calls(b) == 3
    finally:
        N.call("gbr_redwood_destroy", a)
        N.call("gbr_redwood_destroy", b)
