"""Chunk culling (frustum test, two-phase depth-pyramid occlusion test) never
changes a pixel: every plane of every frame is bit-identical to the unculled
renderer and to the oracle, over camera walks, near/far-plane crossings,
narrow fields of view and cameras inside objects."""
import math
import os

import numpy as np
import pytest

from habitat_sim_amd import RendererConfiguration, RendererStandalone, scenes
from habitat_sim_amd import _native as N
from habitat_sim_amd import math3d as m3
from oracle.oracle import OracleRenderer

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)
PLANES = ("color", "depth", "linear_depth", "semantic")
MODES = [N.CULL_FRUSTUM, N.CULL_HIZ]


def _renderer(spec, mode):
    gx, gy = spec.grid
    r = RendererStandalone(RendererConfiguration().set_tile_size_count(spec.tile, (gx, gy)))
    r.set_tuning(cull_mode=mode)
    scenes.populate(r, spec)
    return r


def _same(a, b, what=PLANES, tag=""):
    for name in what:
        x = getattr(a, f"{name}_image")()
        y = getattr(b, f"{name}_image")()
        diff = x != y
        if diff.ndim == 3:
            diff = diff.any(-1)
        assert not diff.any(), f"{tag} {name}: {int(diff.sum())} pixels differ"


def _walk(spec, steps, seed):
    """Seeded random walk per env: turn +-10 degrees or step 0.25 m."""
    rng = np.random.default_rng(seed)
    poses = [env.camera_pose.copy() for env in spec.envs]
    out = []
    for _ in range(steps):
        views = []
        for e in range(len(poses)):
            a = rng.integers(0, 3)
            p = poses[e]
            if a == 1:
                p = m3.mul(p, m3.rotation_y(math.radians(10)))
            elif a == 2:
                p = m3.mul(p, m3.rotation_y(math.radians(-10)))
            else:
                p = p.copy()
                p[3, :3] = p[3, :3] - 0.25 * p[2, :3]
            poses[e] = p
            views.append(m3.inverted_rigid(p))
        out.append(views)
    return out


@pytest.mark.parametrize("mode,tuning", [
    (N.CULL_FRUSTUM, {}), (N.CULL_HIZ, {}), (N.CULL_HIZ, dict(raster_mode=1)),
    # bins overflow: the exact rescan also builds the depth pyramid
    (N.CULL_HIZ, dict(min_bin_capacity=64, bin_capacity_factor=0.02)),
    (N.CULL_HIZ, dict(bin_size=16)), (N.CULL_HIZ, dict(lane_pixels=1)),
    (N.CULL_HIZ, dict(lane_pixels=1024))])
def test_apartment_walk_matches_unculled(mode, tuning):
    spec = scenes.apartments(n_envs=48)
    plain = _renderer(spec, N.CULL_NONE)
    culled = _renderer(spec, mode)
    culled.set_tuning(**tuning)
    walk = _walk(spec, 6, seed=11)
    total = skipped = 0
    for k, views in enumerate(walk):
        for r in (plain, culled):
            for e, v in enumerate(views):
                r.update_camera(e, spec.envs[e].projection, v)
            r.draw()
        _same(plain, culled, tag=f"frame {k}")
        st = culled.frame_stats()
        total += st["work_items"]
        skipped += st["work_items"] - st["items_phase1"] - st["items_phase2"]
        if "bin_capacity_factor" in tuning:
            assert st["overflowed_bins"] > 0
    # culling is actually exercised
    assert skipped > 0.3 * total
    plain.close()
    culled.close()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("far,hfov", [(1000.0, 90.0), (3.0, 90.0), (1000.0, 20.0)])
def test_small_rooms_cameras_vs_oracle(mode, far, hfov):
    """Random cameras, some inside the furniture or against a wall (near
    plane through geometry), far plane through the room, narrow views."""
    spec = scenes.small_rooms(n_envs=16)
    gx, gy = spec.grid
    r = _renderer(spec, mode)
    o = OracleRenderer(spec.tile, (gx, gy))
    scenes.populate(o, spec)
    P = m3.pinhole_projection(spec.tile, math.radians(hfov), 0.01, far)
    rng = np.random.default_rng(5)
    for frame in range(3):
        for e in range(len(spec.envs)):
            pos = (rng.uniform(-3.9, 3.9), rng.uniform(0.05, 2.9), rng.uniform(-2.9, 2.9))
            pose = m3.camera_pose(pos, rng.uniform(0, 2 * math.pi), rng.uniform(-0.6, 0.6))
            for x in (r, o):
                x.update_camera(e, P, m3.inverted_rigid(pose))
        r.draw()
        r.draw()  # second draw: previous-frame proposal populated
        o.draw(NTHREADS)
        _same(r, o, tag=f"frame {frame}")
    r.close()


def test_switching_modes_keeps_output():
    spec = scenes.small_rooms(n_envs=8)
    r = _renderer(spec, N.CULL_NONE)
    r.draw()
    ref = {n: getattr(r, f"{n}_image")() for n in PLANES}
    for mode in (N.CULL_HIZ, N.CULL_FRUSTUM, N.CULL_HIZ, N.CULL_NONE):
        r.set_tuning(cull_mode=mode)
        for _ in range(2):
            r.draw()
            for n in PLANES:
                np.testing.assert_array_equal(getattr(r, f"{n}_image")(), ref[n], err_msg=f"{mode} {n}")
    r.close()


def test_learned_mode_needs_a_model():
    spec = scenes.small_rooms(n_envs=1)
    r = _renderer(spec, N.CULL_NONE)
    with pytest.raises(N.GbrError) as e:
        r.set_tuning(cull_mode=N.CULL_LEARNED)
    assert e.value.status == N.ERR_STATE
    r.close()


def test_visible_items_agree_across_modes_and_recall():
    """The work items owning a visible pixel are the same in every mode (the
    pixels are); phase 1 holds all of them without culling and in frustum
    mode, and most of them with the previous-frame proposal."""
    spec = scenes.apartments(n_envs=16)
    walk = _walk(spec, 3, seed=5)
    vis = {}
    for mode in (N.CULL_NONE, N.CULL_FRUSTUM, N.CULL_HIZ):
        r = _renderer(spec, mode)
        for views in walk:
            for e, v in enumerate(views):
                r.update_camera(e, spec.envs[e].projection, v)
            r.draw()
        vis[mode] = r.frame_visibility()
        r.close()
    n = vis[N.CULL_NONE]["visible_items"]
    assert n > 0
    assert all(v["visible_items"] == n for v in vis.values()), vis
    assert vis[N.CULL_NONE]["recall"] == 1.0 and vis[N.CULL_FRUSTUM]["recall"] == 1.0
    assert 0.5 < vis[N.CULL_HIZ]["recall"] <= 1.0, vis
