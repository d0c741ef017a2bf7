"""Texture-LOD scenes whose GL result has a closed form (GL 4.6 section
8.14): a full-tile quad under an identity camera with an affine texture
mapping, so every pixel quad has the same derivatives, and textures whose
mip levels are each one constant colour, so filtering inside a level returns
that colour and only the level choice and the blend between levels show.

lambda = log2(rho), rho = max(|d(u,v)/dx|, |d(u,v)/dy|) in texels (eq. 8.9);
*_MIPMAP_NEAREST: level d = 0 if lambda <= 1/2, else ceil(lambda + 1/2) - 1
(eq. 8.10); *_MIPMAP_LINEAR: levels floor(lambda) and floor(lambda) + 1,
blended (1 - frac(lambda)) * tau1 + frac(lambda) * tau2 (eq. 8.11).  With
magnification LINEAR and minification NEAREST_MIPMAP_*, lambda <= 1/2 counts
as magnification (section 8.14.2)."""
from __future__ import annotations

import math

import numpy as np

import synth
from habitat_sim_amd.assets import (GL_CLAMP_TO_EDGE, GL_LINEAR, GL_LINEAR_MIPMAP_LINEAR,
                                    GL_LINEAR_MIPMAP_NEAREST, GL_NEAREST_MIPMAP_LINEAR,
                                    GL_REPEAT)

TILE = 8
SIZE = 512  # level 0 edge; 10 levels down to 1x1
# per-level constant colours, red and green from unrelated sequences so that
# neither a swapped blend nor a neighbouring level can match by accident
RED = [0, 240, 40, 200, 80, 160, 120, 20, 250, 60]
GREEN = [255, 10, 180, 90, 30, 220, 140, 70, 5, 110]


def level_texture(min_filter, mag_filter=GL_LINEAR, wrap=GL_REPEAT):
    levels = []
    s, lv = SIZE, 0
    while s >= 1:
        a = np.zeros((s, s, 4), np.uint8)
        a[..., 0] = RED[lv]
        a[..., 1] = GREEN[lv]
        a[..., 3] = 255
        levels.append(a)
        s //= 2
        lv += 1
    return synth.texture(levels, min_filter=min_filter, mag_filter=mag_filter, wrap=wrap)


def quad_mesh(ux, uy, vx, vy, u0=0.0, v0=0.0):
    """Full-tile quad (identity camera) whose texture coordinates are the
    affine map (u, v) = (u0 + ux*X + uy*Y, v0 + vx*X + vy*Y) of window
    position (X, Y) in pixels."""
    corners = [(-1.0, -1.0), (1.0, -1.0), (1.0, 1.0), (-1.0, 1.0)]
    pos, uv = [], []
    for nx, ny in corners:
        X = (nx + 1.0) * TILE / 2
        Y = (ny + 1.0) * TILE / 2
        pos.append((nx, ny, 0.0))
        uv.append((u0 + ux * X + uy * Y, v0 + vx * X + vy * Y))
    return (np.array(pos, np.float32), np.array([0, 1, 2, 0, 2, 3], np.uint32),
            np.array(uv, np.float32), None, 0)


def isotropic(lam):
    """A mapping with rho = 2^lam texels per pixel along both axes."""
    a = 2.0 ** lam / SIZE
    return quad_mesh(a, 0.0, 0.0, a)


def scaled(lam_x, lam_y):
    """rho_x = 2^lam_x (texels per pixel along x), rho_y = 2^lam_y along y:
    lambda = max of the two."""
    return quad_mesh(2.0 ** lam_x / SIZE, 0.0, 0.0, 2.0 ** lam_y / SIZE)


def rotated(lam, theta):
    """rho = 2^lam along both axes, the texture rotated by theta against the
    screen: every derivative has a u and a v part."""
    a = 2.0 ** lam / SIZE
    c, s = math.cos(theta), math.sin(theta)
    return quad_mesh(a * c, -a * s, a * s, a * c, 0.25, 0.25)


def expected_linear(lam):
    """*_MIPMAP_LINEAR colour (red, green) in [0, 255] units, unrounded."""
    d = math.floor(lam)
    f = lam - d
    return tuple((1.0 - f) * seq[d] + f * seq[d + 1] for seq in (RED, GREEN))


def expected_nearest(lam):
    d = 0 if lam <= 0.5 else math.ceil(lam + 0.5) - 1
    return (RED[d], GREEN[d])


def scene(case):
    """(tile, meshes, materials, textures, expected (red, green), tolerance)
    for a named case; tolerance in 8-bit steps (1: the blend of two unorm8
    values rounds once)."""
    kind, args = case
    if kind == "lml":  # LINEAR_MIPMAP_LINEAR at lambda = k + f
        lam = args
        return (TILE, [isotropic(lam)], [synth.flat((1, 1, 1, 1), texture=0)],
                [level_texture(GL_LINEAR_MIPMAP_LINEAR)], expected_linear(lam), 1)
    if kind == "lmn":  # LINEAR_MIPMAP_NEAREST
        lam = args
        return (TILE, [isotropic(lam)], [synth.flat((1, 1, 1, 1), texture=0)],
                [level_texture(GL_LINEAR_MIPMAP_NEAREST)], expected_nearest(lam), 0)
    if kind == "nml":  # NEAREST_MIPMAP_LINEAR (magnification LINEAR: c = 1/2)
        lam = args
        # lambda <= c magnifies: level 0, whatever the minification filter
        exp = (RED[0], GREEN[0]) if lam <= 0.5 else expected_linear(lam)
        return (TILE, [isotropic(lam)], [synth.flat((1, 1, 1, 1), texture=0)],
                [level_texture(GL_NEAREST_MIPMAP_LINEAR)], exp, 1 if lam > 0.5 else 0)
    if kind == "aniso":  # rho_x != rho_y, lambda = log2 of the larger
        lx, ly = args
        return (TILE, [scaled(lx, ly)], [synth.flat((1, 1, 1, 1), texture=0)],
                [level_texture(GL_LINEAR_MIPMAP_LINEAR)], expected_linear(max(lx, ly)), 1)
    if kind == "rot":  # rotated mapping
        lam, theta = args
        return (TILE, [rotated(lam, theta)], [synth.flat((1, 1, 1, 1), texture=0)],
                [level_texture(GL_LINEAR_MIPMAP_LINEAR)], expected_linear(lam), 1)
    raise ValueError(kind)


CASES = ([("lml", k + f) for k in (0, 1, 2) for f in (0.25, 0.5, 0.75)]
         + [("lmn", k + d) for k in (1, 2, 3) for d in (-0.4, 0.4)]
         + [("nml", k + d) for k in (1, 2) for d in (-0.4, 0.4)] + [("nml", 0.3)]
         + [("aniso", (1.3, 2.6)), ("aniso", (2.45, 0.7))]
         + [("rot", (1.7, math.radians(30))), ("rot", (2.3, math.radians(-55)))])


def case_id(case):
    kind, args = case
    if isinstance(args, tuple):
        return kind + "-" + "-".join(f"{a:.3g}" for a in args)
    return f"{kind}-{args:.3g}"


def check(color, expected, tol):
    """Every pixel's red and green within tol of the closed form (rounded to
    the nearest 8-bit value first)."""
    red = np.asarray(color)[..., 0].astype(np.int64)
    green = np.asarray(color)[..., 1].astype(np.int64)
    er, eg = (int(math.floor(e + 0.5)) for e in expected)
    return (np.abs(red - er) <= tol).all() and (np.abs(green - eg) <= tol).all()


__all__ = ["CASES", "case_id", "scene", "check", "level_texture", "GL_CLAMP_TO_EDGE"]
