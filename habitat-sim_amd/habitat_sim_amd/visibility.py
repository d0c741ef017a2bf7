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


def random_poses(rng, n: int, height: float = 1.5, pitch_deg: float = 15.0) -> np.ndarray:
    """n camera-to-world poses (n, 4, 4), uniform position in the apartment
    footprint, uniform yaw, pitch uniform in +-pitch_deg."""
    x = rng.uniform(0.3, ROOM * GRID - 0.3, n)
    z = rng.uniform(0.3, ROOM * GRID - 0.3, n)
    yaw = rng.uniform(0.0, 2 * math.pi, n)
    pitch = np.radians(rng.uniform(-pitch_deg, pitch_deg, n))
    cy, sy = np.cos(yaw).astype(np.float32), np.sin(yaw).astype(np.float32)
    cp, sp = np.cos(pitch).astype(np.float32), np.sin(pitch).astype(np.float32)
    out = np.zeros((n, 4, 4), np.float32)
    # rotation_y(yaw) * rotation_x(pitch), [column][row]
    out[:, 0, 0], out[:, 0, 2] = cy, -sy
    out[:, 1, 0], out[:, 1, 1], out[:, 1, 2] = sy * sp, cp, cy * sp
    out[:, 2, 0], out[:, 2, 1], out[:, 2, 2] = sy * cp, -sp, cy * cp
    out[:, 3, 0], out[:, 3, 1], out[:, 3, 2], out[:, 3, 3] = x, height, z, 1.0
    return out


def inverted_rigid(poses: np.ndarray) -> np.ndarray:
    """View matrices of camera-to-world poses (n, 4, 4): R^T, -R^T t."""
    poses = np.asarray(poses, np.float32)
    out = np.zeros_like(poses)
    r = poses[:, :3, :3]
    out[:, :3, :3] = np.transpose(r, (0, 2, 1))
    out[:, 3, :3] = -np.einsum("ncr,nr->nc", r, poses[:, 3, :3]).astype(np.float32)
    out[:, 3, 3] = 1.0
    return out


def _turn(poses: np.ndarray, rot: np.ndarray) -> np.ndarray:
    """poses * rot for a batch, accumulated like math3d.mul."""
    out = np.zeros_like(poses)
    for k in range(4):
        out = (out + (poses[:, k, None, :] * rot[None, :, k, None]).astype(np.float32)).astype(np.float32)
    return out


def walk_poses(start: np.ndarray, steps: int, rng) -> np.ndarray:
    """The seeded random walk from poses `start` (n, 4, 4): per step and env
    an action drawn uniformly from 0.25 m forward / 10 degrees left / 10
    degrees right, positions clamped to the apartment footprint; returns the
    poses after each step, (steps, n, 4, 4)."""
    poses = np.array(start, np.float32)
    left, right = m3.rotation_y(math.radians(10)), m3.rotation_y(math.radians(-10))
    lo, hi = 0.3, ROOM * GRID - 0.3
    out = np.empty((steps,) + poses.shape, np.float32)
    for s in range(steps):
        acts = rng.integers(0, 3, len(poses))
        nxt = poses.copy()
        for a, rot in ((1, left), (2, right)):
            sel = acts == a
            if sel.any():
                nxt[sel] = _turn(poses[sel], rot)
        fwd = acts == 0
        if fwd.any():
            p = poses[fwd]
            pos = np.clip(p[:, 3, :3] - np.float32(0.25) * p[:, 2, :3], lo, hi).astype(np.float32)
            pos[:, 1] = 1.5
            nxt[fwd, 3, :3] = pos
        poses = nxt
        out[s] = poses
    return out


def camera_walk(spec: BatchSpec, steps: int, seed: int, start=None) -> np.ndarray:
    """Per step the view matrices of every env, (steps, n, 4, 4): from the
    env's own pose (or `start` poses) turn +-10 degrees or move 0.25 m
    forward at random, clamped to the apartment footprint."""
    rng = np.random.default_rng(seed)
    if start is None:
        start = np.stack([e.camera_pose for e in spec.envs])
    poses = walk_poses(np.asarray(start, np.float32), steps, rng)
    return inverted_rigid(poses.reshape(-1, 4, 4)).reshape(poses.shape)


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
        start = random_poses(rng, len(spec.envs))
        if walk_length == 1:  # uniform random poses, no walk
            walk = inverted_rigid(start)[None]
        else:
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
