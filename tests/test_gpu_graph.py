"""Frames replayed from a captured HIP graph are the frames drawn directly:
a camera walk drawn by a renderer with graph replay and by one without, every
plane of every frame compared, with host reads, direct (timed) frames and
transform edits between the replays."""
import numpy as np
import pytest

from habitat_sim_amd import RendererConfiguration, RendererStandalone, scenes
from habitat_sim_amd import _native as N
from habitat_sim_amd import math3d as m3

from test_gpu_cull import _same, _walk

pytestmark = pytest.mark.gpu


def _renderer(spec, mode, graph):
    gx, gy = spec.grid
    r = RendererStandalone(RendererConfiguration().set_tile_size_count(spec.tile, (gx, gy)))
    r.set_tuning(cull_mode=mode, graph=graph)
    scenes.populate(r, spec)
    return r


@pytest.mark.parametrize("mode", [N.CULL_NONE, N.CULL_HIZ])
def test_graph_replay_matches_direct_draws(mode):
    spec = scenes.apartments(n_envs=24, tile=64, instances=60)
    a = _renderer(spec, mode, graph=1)
    b = _renderer(spec, mode, graph=0)
    walk = _walk(spec, 8, seed=11)
    for k, views in enumerate(walk):
        for r in (a, b):
            for e, env in enumerate(spec.envs):
                r.update_camera(e, env.projection, views[e])
        if k == 3:
            # a direct frame in the middle (stage events recorded): the next
            # replay must pick up the device frame counter it advanced
            a.set_tuning(timing=1)
        if k == 4:
            a.set_tuning(timing=0)
        if k == 5:
            # transforms change between replays: uploaded outside the graph
            for r in (a, b):
                t = r.transformations(0)
                t[1] = m3.mul(m3.translation((0.3, 0.0, 0.0)), np.asarray(t[1]))
        a.draw()
        b.draw()
        _same(a, b, tag=f"frame {k}")
    a.close()
    b.close()


def test_graph_is_rebuilt_when_the_scene_changes():
    spec = scenes.small_rooms(n_envs=4)
    a = _renderer(spec, N.CULL_NONE, graph=1)
    b = _renderer(spec, N.CULL_NONE, graph=0)
    for r in (a, b):
        r.draw()
    _same(a, b, tag="before")
    for r in (a, b):
        r.clear(1)
        r.draw()
    _same(a, b, tag="after clear")
    a.close()
    b.close()
