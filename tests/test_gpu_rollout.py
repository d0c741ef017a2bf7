"""The trajectory-rollout workload at full size: 4096 apartment envs (~200k
triangles each), 128x128 linear depth + semantic, 500 steps of the seeded
random walk (0.25 m forward / 10 degrees left / right, clamped to the
apartment) in learned mode with the previous frame's visible chunks, rows
collected every frame and the proposal model fine-tuned every 50 steps.
Every 50th frame equals a plain (unculled) render of the same views bit for
bit; frame latency p50/p99 and the proposal's recall are reported."""
import time

import numpy as np
import pytest

from habitat_sim_amd import RendererConfiguration, RendererStandalone, scenes, visibility
from habitat_sim_amd import _native as N

pytestmark = pytest.mark.gpu

ENVS, STEPS, EVERY = 4096, 500, 50


def test_rollout_4096_envs_500_steps_fine_tuned_stays_exact():
    spec = scenes.apartments(n_envs=ENVS, tile=128)
    gx, gy = spec.grid
    planes = N.PLANE_LINEAR_DEPTH | N.PLANE_SEMANTIC

    def make():
        cfg = RendererConfiguration().set_tile_size_count(spec.tile, (gx, gy))
        r = RendererStandalone(cfg, planes=planes, top_down=True)
        scenes.populate(r, spec)
        return r

    plain, learned = make(), make()
    visibility.train(learned, spec, frames=8, steps=200, seed=3)
    learned.set_tuning(cull_mode=N.CULL_LEARNED, proposal_history=1)
    walk = visibility.camera_walk(spec, STEPS, seed=11)
    P = np.stack([e.projection for e in spec.envs])
    w0 = learned.vis_weights().copy()
    lat, recall, checked = [], [], []

    def check(k):
        if (k + 1) % EVERY:
            return
        plain.update_cameras(0, P, walk[k])
        plain.draw()
        for name in ("linear_depth", "semantic"):
            a = getattr(plain, f"{name}_image")()
            b = getattr(learned, f"{name}_image")()
            assert np.array_equal(a, b), f"step {k + 1}: {name} differs in {int((a != b).sum())} px"
        checked.append(k + 1)

    learned.vis_clear_rows()
    fits = []
    for k in range(STEPS):
        t0 = time.perf_counter()
        learned.update_cameras(0, P, walk[k])
        learned.vis_collect()
        learned.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
        recall.append(learned.frame_visibility()["recall"])
        check(k)
        if (k + 1) % EVERY == 0:
            fits.append(learned.vis_fit(steps=20, batch=4096, learning_rate=1e-3, seed=100 + k))
            learned.vis_clear_rows()
    assert checked == list(range(EVERY, STEPS + 1, EVERY))
    assert len(fits) == STEPS // EVERY and all(np.isfinite(f["loss_last"]) for f in fits)
    assert not np.array_equal(learned.vis_weights(), w0)  # fine-tuning moved the model
    p50, p99 = np.percentile(lat, 50), np.percentile(lat, 99)
    assert np.isfinite(p50) and p99 >= p50 > 0
    assert 0.0 < min(recall) <= max(recall) <= 1.0
    print(f"rollout {ENVS} envs x {STEPS} steps: latency p50 {p50:.2f} ms, p99 {p99:.2f} ms; "
          f"recall mean {np.mean(recall):.3f}, min {min(recall):.3f}")
    plain.close()
    learned.close()
