"""Training the learned visibility proposal of the batch renderer.

The renderer scores every chunk that passes the frustum test with a small
MLP (32 features -> 128 -> 128 -> 1); in cull mode CULL_LEARNED the chunks
scored visible are drawn first and the rest are tested against their depth
pyramid.  The proposal only orders work, so a poor model costs time, never
pixels.  Training rows are (features, chunk had a visible pixel) pairs taken
from frames the renderer draws on camera walks through the scenes; the
network is trained on the GPU by the library's own BF16 MFMA forward and
backward passes with Adam (gbr_vis_* in include/gbr.h).
"""
from __future__ import annotations

import math

import numpy as np

from . import _native as N
from . import math3d as m3
from .scenes import GRID, ROOM, BatchSpec


def random_pose(rng, height: float = 1.5, pitch_deg: float = 15.0) -> np.ndarray:
    """Camera-to-world pose uniformly inside the apartment footprint."""
    x = rng.uniform(0.3, ROOM * GRID - 0.3)
    z = rng.uniform(0.3, ROOM * GRID - 0.3)
    yaw = rng.uniform(0.0, 2 * math.pi)
    pitch = math.radians(rng.uniform(-pitch_deg, pitch_deg))
    return m3.camera_pose((x, height, z), yaw, pitch)


def camera_walk(spec: BatchSpec, steps: int, seed: int, start=None) -> list:
    """Per step a list of view matrices, one per env: from the env's own pose
    (or `start` poses) turn +-10 degrees or move 0.25 m forward at random,
    clamped to the apartment footprint."""
    rng = np.random.default_rng(seed)
    poses = [p.copy() for p in (start if start is not None else [e.camera_pose for e in spec.envs])]
    out = []
    lo, hi = 0.3, ROOM * GRID - 0.3
    for _ in range(steps):
        acts = rng.integers(0, 3, len(poses))
        views = []
        for e, p in enumerate(poses):
            if acts[e] == 1:
                p = m3.mul(p, m3.rotation_y(math.radians(10)))
            elif acts[e] == 2:
                p = m3.mul(p, m3.rotation_y(math.radians(-10)))
            else:
                p = p.copy()
                p[3, :3] = np.clip(p[3, :3] - 0.25 * p[2, :3], lo, hi).astype(np.float32)
                p[3, 1] = 1.5
            poses[e] = p
            views.append(m3.inverted_rigid(p))
        out.append(views)
    return out


def train(renderer, spec: BatchSpec, frames: int = 64, steps: int = 600, seed: int = 1,
          walk_length: int = 16, batch: int = 8192, learning_rate: float = 2e-3,
          pos_weight: float = 2.0) -> dict:
    """Collects rows from `frames` frames (walks of `walk_length` steps from
    random poses, depth-pyramid culling so that the visibility history looks
    as it does when rendering), then fits the model.  Restores the cameras
    and the cull mode; returns the fit report."""
    t = renderer.tuning()
    mode = t.cull_mode
    renderer.vis_init(seed)
    renderer.vis_clear_rows()
    renderer.set_tuning(cull_mode=N.CULL_HIZ)
    rng = np.random.default_rng(seed ^ 0x5A17)
    done = 0
    while done < frames:
        start = [random_pose(rng) for _ in spec.envs]
        walk = camera_walk(spec, min(walk_length, frames - done), int(rng.integers(1 << 30)),
                           start=start)
        for views in walk:
            renderer.update_cameras(0, np.stack([e.projection for e in spec.envs]),
                                    np.stack(views))
            renderer.vis_collect()
            done += 1
    report = renderer.vis_fit(steps=steps, batch=batch, learning_rate=learning_rate,
                              pos_weight=pos_weight, seed=seed)
    for e, env in enumerate(spec.envs):
        renderer.update_camera(e, env.projection, m3.inverted_rigid(env.camera_pose))
    renderer.set_tuning(cull_mode=mode)
    return report


def online_walk(renderer, spec: BatchSpec, walk: list, finetune_every: int = 50,
                fit_steps: int = 20, batch: int = 4096, learning_rate: float = 1e-3,
                seed: int = 1, on_frame=None) -> list:
    """Renders a camera walk (per step one view matrix per env) in the
    renderer's current cull mode while fine-tuning the visibility model
    online: every frame also appends its training rows, and every
    `finetune_every` frames the model takes `fit_steps` Adam steps on the
    rows gathered since the previous fine-tune (then they are dropped).
    on_frame(step) is called after each frame; returns the fit reports."""
    reports = []
    projections = np.stack([e.projection for e in spec.envs])
    renderer.vis_clear_rows()
    for k, views in enumerate(walk):
        renderer.update_cameras(0, projections, np.stack(views))
        renderer.vis_collect()
        if on_frame is not None:
            on_frame(k)
        if (k + 1) % finetune_every == 0:
            reports.append(renderer.vis_fit(steps=fit_steps, batch=batch,
                                            learning_rate=learning_rate, seed=seed + k))
            renderer.vis_clear_rows()
    return reports
