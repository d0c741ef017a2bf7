"""The autotuner times candidate launch parameters and leaves the fastest;
every candidate renders the same pixels, so the planes after tuning equal
those before."""
import numpy as np
import pytest

from habitat_sim_amd import RendererConfiguration, RendererStandalone, scenes
from habitat_sim_amd import _native as N

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [N.CULL_NONE, N.CULL_HIZ])
def test_autotune_keeps_pixels_and_sets_the_fastest(mode):
    spec = scenes.apartments(n_envs=16, tile=64, instances=40)
    gx, gy = spec.grid
    r = RendererStandalone(RendererConfiguration().set_tile_size_count(spec.tile, (gx, gy)))
    r.set_tuning(cull_mode=mode)
    scenes.populate(r, spec)
    r.draw()
    before = {k: getattr(r, f"{k}_image")() for k in ("color", "depth", "semantic")}
    rep = r.autotune(warmup=1, frames=2)
    assert rep["candidates"] >= 10
    assert 0.0 < rep["best_ms"] <= rep["initial_ms"]
    t = r.tuning()
    for k, v in rep["best"].items():
        assert getattr(t, k) == v, k
    assert t.cull_mode == mode
    r.draw()
    for k, v in before.items():
        assert np.array_equal(getattr(r, f"{k}_image")(), v), k
    r.close()


def test_autotune_rejects_bad_parameters():
    spec = scenes.small_rooms(n_envs=1)
    r = RendererStandalone(RendererConfiguration().set_tile_size_count(spec.tile, (1, 1)))
    with pytest.raises(Exception):
        r.autotune(frames=0)
    r.close()
