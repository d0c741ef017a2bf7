"""Matrix helpers with Magnum's conventions: 4x4 float32 arrays indexed
[column][row] (column-major, the layout Renderer::transformations() and
updateCamera() take), built the way Magnum's Matrix4 factory functions build
them (float32 arithmetic).

Cited formulas:
  - Magnum Matrix4::orthographicProjection / perspectiveProjection, used by
    src/esp/sensor/CameraSensor.cpp:45-56 and src/tests/GfxBatchRendererTest.cpp
  - CameraSensorSpec::projectionMatrix(): pinhole P from hfov, near-plane
    width 2*near*tan(hfov/2), height = width * rows/cols
    (src/esp/sensor/CameraSensor.cpp:45-66; defaults near 0.01, far 1000,
    hfov 90 deg, src/esp/sensor/VisualSensor.h:52,56,225)
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32


def identity() -> np.ndarray:
    return np.eye(4, dtype=np.float32)


def translation(v) -> np.ndarray:
    m = identity()
    m[3, :3] = np.asarray(v, np.float32)
    return m


def scaling(v) -> np.ndarray:
    v = np.broadcast_to(np.asarray(v, np.float32), (3,))
    m = identity()
    m[0, 0], m[1, 1], m[2, 2] = v
    return m


def rotation_x(rad: float) -> np.ndarray:
    c, s = f32(math.cos(rad)), f32(math.sin(rad))
    m = identity()
    m[1, 1], m[1, 2], m[2, 1], m[2, 2] = c, s, -s, c
    return m


def rotation_y(rad: float) -> np.ndarray:
    c, s = f32(math.cos(rad)), f32(math.sin(rad))
    m = identity()
    m[0, 0], m[0, 2], m[2, 0], m[2, 2] = c, -s, s, c
    return m


def rotation_z(rad: float) -> np.ndarray:
    c, s = f32(math.cos(rad)), f32(math.sin(rad))
    m = identity()
    m[0, 0], m[0, 1], m[1, 0], m[1, 1] = c, s, -s, c
    return m


def mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Matrix product a*b on [col][row] arrays, accumulated in float32 over
    k = 0..3 from +0 (Magnum's RectangularMatrix::operator*)."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    out = np.zeros((4, 4), np.float32)
    for k in range(4):
        # out[c][r] += a[k][r] * b[c][k]
        out = (out + (a[k][None, :] * b[:, k][:, None]).astype(np.float32)).astype(np.float32)
    return out


def inverted(m: np.ndarray) -> np.ndarray:
    """General inverse (computed in float64, rounded to float32)."""
    return np.linalg.inv(np.asarray(m, np.float64).T).T.astype(np.float32)


def inverted_rigid(m: np.ndarray) -> np.ndarray:
    """Inverse of a rotation+translation matrix: R^T, -R^T t."""
    m = np.asarray(m, np.float32)
    out = identity()
    r = m[:3, :3]            # [col][row]
    out[:3, :3] = r.T
    t = m[3, :3]
    out[3, :3] = -(r @ t).astype(np.float32)
    return out


def orthographic_projection(size, near: float, far: float) -> np.ndarray:
    size = np.asarray(size, np.float32)
    xy = (f32(2.0) / size).astype(np.float32)
    z = f32(2.0) / (f32(near) - f32(far))
    m = np.zeros((4, 4), np.float32)
    m[0, 0], m[1, 1], m[2, 2] = xy[0], xy[1], z
    m[3, 2] = f32(f32(near) * z) - f32(1.0)
    m[3, 3] = 1.0
    return m


def perspective_projection_size(size, near: float, far: float) -> np.ndarray:
    size = np.asarray(size, np.float32)
    near, far = f32(near), f32(far)
    xy = (f32(2.0) * near / size).astype(np.float32)
    z = f32(1.0) / (near - far)
    m = np.zeros((4, 4), np.float32)
    m[0, 0], m[1, 1] = xy[0], xy[1]
    m[2, 2] = f32((far + near) * z)
    m[2, 3] = -1.0
    m[3, 2] = f32(f32(f32(2.0) * far) * near) * z
    return m


def perspective_projection(fov_rad: float, aspect: float, near: float, far: float) -> np.ndarray:
    xy = f32(f32(2.0) * f32(math.tan(f32(fov_rad) / f32(2.0)))) * f32(near)
    return perspective_projection_size((xy, xy / f32(aspect)), near, far)


def pinhole_projection(resolution_hw, hfov_rad: float = math.pi / 2, near: float = 0.01,
                       far: float = 1000.0) -> np.ndarray:
    """CameraSensorSpec::projectionMatrix() for a pinhole sensor."""
    h, w = resolution_hw
    near_plane = np.array([1.0, f32(h) / f32(w)], np.float32)
    scale = f32(1.0) / (f32(2.0) * f32(near) * f32(math.tan(f32(0.5) * f32(hfov_rad))))
    return perspective_projection_size(near_plane / scale, near, far)


def camera_pose(position, yaw: float, pitch: float) -> np.ndarray:
    """Camera-to-world transform: yaw about +Y, then pitch about +X, camera
    looking down -Z (habitat convention, Y up)."""
    m = mul(rotation_y(yaw), rotation_x(pitch))
    m[3, :3] = np.asarray(position, np.float32)
    return m


def look_at(eye, target, up) -> np.ndarray:
    """Matrix4::lookAt(eye, target, up): camera-style transformation whose
    -Z looks from eye at target (columns right, up, backward, eye)."""
    eye = np.asarray(eye, np.float32)
    backward = eye - np.asarray(target, np.float32)
    backward = (backward / np.float32(np.sqrt(np.dot(backward, backward)))).astype(np.float32)
    right = np.cross(np.asarray(up, np.float32), backward).astype(np.float32)
    right = (right / np.float32(np.sqrt(np.dot(right, right)))).astype(np.float32)
    real_up = np.cross(backward, right).astype(np.float32)
    m = np.zeros((4, 4), np.float32)
    m[0, :3], m[1, :3], m[2, :3], m[3, :3] = right, real_up, backward, eye
    m[3, 3] = 1.0
    return m
