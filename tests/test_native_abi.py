"""CPU-side checks of the C ABI library: it loads, exports every entry point
include/gbr.h declares, and its host-only entry points behave like the
reference (no GPU needed)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from habitat_sim_amd import _native as N
from habitat_sim_amd import calculate_depth_unprojection, math3d as m3, unproject_depth
from oracle.oracle import depth_unprojection as orc_depth_unprojection
from oracle.oracle import unproject_depth as orc_unproject_depth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "gbr.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gbr_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = N.lib()
    declared = _declared()
    assert len(declared) >= 35
    for name in declared:
        assert hasattr(lib, name), f"{name} not exported by libgbr.so"
        assert name in N.SIGNATURES, f"{name} has no ctypes signature"
    assert set(N.SIGNATURES) == set(declared)


def test_abi_version_and_config_defaults():
    lib = N.lib()
    assert lib.gbr_abi_version() == 1
    c = N.Config()
    N.check(lib.gbr_config_init(ctypes.byref(c)))
    # RendererConfiguration defaults, Renderer.cpp:58-64
    assert tuple(c.tile_size) == (128, 128) and tuple(c.tile_count) == (1, 1)
    assert c.flags == 0 and c.max_light_count == 0
    assert abs(c.ambient_factor - 0.1) < 1e-7
    assert c.planes == N.PLANE_ALL


def test_tuning_defaults():
    lib = N.lib()
    t = N.Tuning()
    N.check(lib.gbr_tuning_init(ctypes.byref(t)))
    assert (t.bin_size, t.chunk_triangles, t.queue_grain) == (32, 1024, 4)
    assert (t.cull_mode, t.raster_mode, t.timing) == (N.CULL_NONE, 0, 0)
    assert t.lane_pixels == 64 and t.proposal_threshold == 0.0 and t.graph == 1
    assert t.heavy_entries == 256 and t.proposal_history == 0


def test_null_arguments_are_errors_not_crashes():
    lib = N.lib()
    assert lib.gbr_create(None, None) == N.ERR_ARGUMENT
    assert b"null" in lib.gbr_last_error()
    assert lib.gbr_draw(None) == N.ERR_ARGUMENT
    assert lib.gbr_destroy(None) == N.OK


@pytest.mark.parametrize("near,far", [(0.01, 100.0), (1.0, 100.0), (0.001, 10.0), (0.1, 1000.0)])
def test_depth_unprojection_matches_oracle_bitwise(near, far):
    P = m3.perspective_projection(math.radians(60.0), 1.0, near, far)
    a = calculate_depth_unprojection(P)
    b = orc_depth_unprojection(P)
    assert a.tobytes() == b.tobytes()
    rng = np.random.default_rng(0)
    d = rng.uniform(0, 1, (33, 17)).astype(np.float32)
    d[0, :5] = 1.0
    got = unproject_depth(a, d.copy())
    exp = orc_unproject_depth(b, d).reshape(d.shape)
    assert got.tobytes() == exp.tobytes()
    assert (got[0, :5] == 0.0).all()  # far plane -> 0


def test_create_reports_device_state():
    """Without a GPU gbr_create fails with GBR_ERR_DEVICE (no crash); with one
    it succeeds and destroys cleanly."""
    lib = N.lib()
    c = N.Config()
    lib.gbr_config_init(ctypes.byref(c))
    h = ctypes.c_void_p()
    st = lib.gbr_create(ctypes.byref(c), ctypes.byref(h))
    if st == N.OK:
        assert lib.gbr_destroy(h) == N.OK
    else:
        assert st == N.ERR_DEVICE


def test_invalid_config_rejected_before_device_use():
    lib = N.lib()
    c = N.Config()
    lib.gbr_config_init(ctypes.byref(c))
    c.tile_size[0] = 0
    h = ctypes.c_void_p()
    assert lib.gbr_create(ctypes.byref(c), ctypes.byref(h)) == N.ERR_ARGUMENT
    lib.gbr_config_init(ctypes.byref(c))
    c.max_light_count = N.MAX_LIGHTS + 1
    assert lib.gbr_create(ctypes.byref(c), ctypes.byref(h)) == N.ERR_UNSUPPORTED
