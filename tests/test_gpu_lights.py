"""Phong lights through the C ABI (addLight(), lightColors(), lightRanges(),
clearLights(), Renderer.cpp:990-1121): lit batches bit-exact with the
oracle over every plane, with culling, moving lights, light colours and
ranges edited between frames, more lights than the maximum, and the
reference's assertions as error codes."""
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
MAX_LIGHTS = 4
PLANES = ("color", "depth", "linear_depth", "semantic")


def _shaded(spec):
    """The synthetic scenes are flat-shaded: make every material shaded and
    give every mesh smooth normals (away from its centroid, as on a sphere)."""
    for m in spec.file.materials:
        m.flat = False
    for m in spec.file.meshes:
        p = np.asarray(m.positions, np.float64)
        n = p - p.mean(0)
        n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-6)
        m.normals = n.astype(np.float32)
    return spec


def _add_lights(x, spec, seed):
    """Per env up to MAX_LIGHTS + 1 lights on new empty nodes: point lights
    with finite and infinite ranges, directional lights; env 0 has none."""
    rng = np.random.default_rng(seed)
    for e in range(len(spec.envs)):
        count = 0 if e == 0 else (MAX_LIGHTS + 1 if e == 1 else int(rng.integers(1, MAX_LIGHTS + 1)))
        for i in range(count):
            node = x.add_empty_node(e)
            t = x.transformations(e)
            if i % 3 == 2:
                t[node] = m3.look_at((0, 0, 0), tuple(rng.uniform(-1, 1, 3)), (0.0, 1.0, 0.0))
                kind = N.LIGHT_DIRECTIONAL
            else:
                t[node] = m3.translation(tuple(rng.uniform((-3.5, 0.2, -2.5), (3.5, 2.8, 2.5))))
                kind = N.LIGHT_POINT
            assert x.add_light(e, node, kind) == i
            x.light_colors(e)[i] = rng.uniform(0.0, 2.0, 3).astype(np.float32)
            x.light_ranges(e)[i] = np.inf if i % 2 else np.float32(rng.uniform(1.0, 6.0))


def _pair(spec, tuning=None):
    gx, gy = spec.grid
    r = RendererStandalone(RendererConfiguration().set_tile_size_count(spec.tile, (gx, gy))
                           .set_max_light_count(MAX_LIGHTS))
    if tuning:
        r.set_tuning(**tuning)
    o = OracleRenderer(spec.tile, (gx, gy), max_light_count=MAX_LIGHTS)
    for x in (r, o):
        scenes.populate(x, spec)
        _add_lights(x, spec, seed=7)
    return r, o


def _same(r, o, tag=""):
    for name in PLANES:
        a = getattr(r, f"{name}_image")()
        b = getattr(o, f"{name}_image")()
        diff = a != b
        if diff.ndim == 3:
            diff = diff.any(-1)
        assert not diff.any(), f"{tag} {name}: {int(diff.sum())} pixels differ from the oracle"


@pytest.mark.parametrize("tuning", [None, dict(cull_mode=N.CULL_HIZ), dict(bin_size=16, raster_mode=1)])
def test_lit_small_rooms_bit_exact(tuning):
    spec = _shaded(scenes.small_rooms(n_envs=8))
    r, o = _pair(spec, tuning)
    r.draw()
    o.draw(NTHREADS)
    _same(r, o)
    # lighting actually changes colours: env 0 (no lights) is ambient only
    col = r.color_image()
    assert col[..., :3].max() > 60
    r.close()


def test_lit_apartments_textured_bit_exact():
    spec = _shaded(scenes.apartments(n_envs=4))
    r, o = _pair(spec)
    r.draw()
    o.draw(NTHREADS)
    _same(r, o)
    r.close()


def test_lights_follow_nodes_colours_and_ranges_between_frames():
    spec = _shaded(scenes.small_rooms(n_envs=6))
    r, o = _pair(spec)
    r.draw()
    o.draw(NTHREADS)
    _same(r, o, "frame 0")
    first = r.color_image()
    rng = np.random.default_rng(3)
    for frame in range(1, 4):
        for x in (r, o):
            for e in range(1, len(spec.envs)):
                t = x.transformations(e)
                n = len(t) - 1  # the last light's node
                t[n] = m3.mul(m3.translation((0.25 * frame, 0.0, -0.1 * frame)), np.asarray(t[n]))
        cols = rng.uniform(0.0, 1.5, 3).astype(np.float32)
        rgs = np.float32(rng.uniform(1.0, 4.0))
        for x in (r, o):
            x.light_colors(2)[0] = cols
            x.light_ranges(2)[0] = rgs
        if frame == 3:
            for x in (r, o):
                x.clear_lights(3)
        r.draw()
        o.draw(NTHREADS)
        _same(r, o, f"frame {frame}")
    assert not np.array_equal(r.color_image(), first)
    r.close()


def test_light_api_errors_and_defaults():
    cfg = RendererConfiguration().set_tile_size_count((32, 32), (2, 1))
    r = RendererStandalone(cfg)
    r.add_empty_node(0)
    with pytest.raises(N.GbrError) as e:
        r.add_light(0, 0, N.LIGHT_POINT)
    assert e.value.status == N.ERR_STATE  # max light count is zero
    r.close()
    with pytest.raises(N.GbrError) as e:
        RendererStandalone(RendererConfiguration().set_max_light_count(N.MAX_LIGHTS + 1))
    assert e.value.status == N.ERR_UNSUPPORTED
    r = RendererStandalone(cfg.set_max_light_count(2))
    assert r.max_light_count == 2
    node = r.add_empty_node(1)
    with pytest.raises(N.GbrError) as e:
        r.add_light(1, node + 1, N.LIGHT_POINT)
    assert e.value.status == N.ERR_RANGE
    with pytest.raises(N.GbrError) as e:
        r.add_light(2, 0, N.LIGHT_POINT)
    assert e.value.status == N.ERR_RANGE
    with pytest.raises(N.GbrError) as e:
        r.add_light(1, node, 7)
    assert e.value.status == N.ERR_ARGUMENT
    for fn in (r.light_colors, r.light_ranges):
        with pytest.raises(N.GbrError) as e:
            fn(2)
        assert e.value.status == N.ERR_RANGE
    assert r.add_light(1, node, N.LIGHT_DIRECTIONAL) == 0
    assert r.add_light(1, node, N.LIGHT_POINT) == 1
    np.testing.assert_array_equal(r.light_colors(1), np.ones((2, 3), np.float32))
    assert np.isinf(r.light_ranges(1)).all()
    assert len(r.light_colors(0)) == 0
    r.light_colors(1)[1] = (0.5, 0.25, 0.125)
    r.light_ranges(1)[0] = 3.0
    np.testing.assert_array_equal(r.light_colors(1)[1], np.array([0.5, 0.25, 0.125], np.float32))
    assert r.light_ranges(1)[0] == 3.0 and np.isinf(r.light_ranges(1)[1])
    r.clear_lights(1)
    assert len(r.light_colors(1)) == 0 and len(r.transformations(1)) == 1
    r.add_light(1, node, N.LIGHT_POINT)
    r.clear(1)
    assert len(r.light_colors(1)) == 0 and len(r.transformations(1)) == 0
    r.close()


def test_lit_uneven_batch_matches_the_oracle():
    """Lights in an uneven batch: per-scene viewports and sensor sets, the
    lit resolve bit-exact with the oracle on every produced plane."""
    spec = _shaded(scenes.heterogeneous(n_envs=12, tri_range=(2e4, 2e5), resolutions=(32, 64),
                                        res_p=(0.5, 0.5)))
    gx, gy = spec.grid
    r = RendererStandalone(RendererConfiguration().set_tile_size_count(spec.tile, (gx, gy))
                           .set_max_light_count(MAX_LIGHTS))
    o = OracleRenderer(spec.tile, (gx, gy), max_light_count=MAX_LIGHTS)
    for x in (r, o):
        scenes.populate(x, spec)
        _add_lights(x, spec, seed=21)
    r.draw()
    o.draw(NTHREADS)
    bits = {"color": N.PLANE_COLOR, "depth": N.PLANE_DEPTH, "linear_depth": N.PLANE_LINEAR_DEPTH,
            "semantic": N.PLANE_SEMANTIC}
    tw, th = spec.tile
    for name, bit in bits.items():
        a = getattr(r, f"{name}_image")()
        b = getattr(o, f"{name}_image")()
        for s, env in enumerate(spec.envs):
            if env.planes and not env.planes & bit:
                continue
            y, x = divmod(s, gx)
            d = a[y * th:(y + 1) * th, x * tw:(x + 1) * tw] != b[y * th:(y + 1) * th, x * tw:(x + 1) * tw]
            if d.ndim == 3:
                d = d.any(-1)
            assert not d.any(), f"scene {s} {name}: {int(d.sum())} pixels differ"
    r.close()


def test_ambient_factor_and_no_textures_flag():
    """setAmbientFactor (Renderer.h:107) scales the shaded materials' ambient
    term; NoTextures drops texture sampling from both terms."""
    from habitat_sim_amd import _native as NN
    spec = _shaded(scenes.apartments(n_envs=2))
    gx, gy = spec.grid
    for factor, flags in ((0.35, 0), (0.1, NN.FLAG_NO_TEXTURES)):
        cfg = (RendererConfiguration().set_tile_size_count(spec.tile, (gx, gy))
               .set_max_light_count(MAX_LIGHTS).set_ambient_factor(factor).set_flags(flags))
        r = RendererStandalone(cfg)
        o = OracleRenderer(spec.tile, (gx, gy), flags, max_light_count=MAX_LIGHTS,
                           ambient_factor=factor)
        for x in (r, o):
            scenes.populate(x, spec)
            _add_lights(x, spec, seed=5)
        r.draw()
        o.draw(NTHREADS)
        _same(r, o, f"factor {factor} flags {flags}")
        r.close()
