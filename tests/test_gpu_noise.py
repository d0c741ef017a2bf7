"""The Redwood depth noise kernel (libgbr.so, gbr_redwood_*) against the
oracle: bit-exact with caller-given deviates over batches and ragged sizes,
the generator's Philox blocks bit-exact, zero noise exact, and the
reference's own check (GfxBatch-independent tests/test_noise_models.py:
mean of 100 GPU simulations vs mean of 100 CPU ones on a 0-20 m ramp)."""
import os

import numpy as np
import pytest

from habitat_sim_amd import _native as N
from habitat_sim_amd import hip
from habitat_sim_amd.noise import RedwoodDepthNoiseModel, load_model
from oracle.oracle import philox4x32, redwood_noise

pytestmark = pytest.mark.gpu

MODEL = os.path.join(os.path.dirname(__file__), "golden", "noise", "redwood-depth-dist-model.npy")


def _host(buf, shape, dtype=np.float32):
    out = np.empty(shape, dtype)
    hip.hip().hipDeviceSynchronize()
    hip.memcpy_d2h(out, buf.value)
    return out


@pytest.mark.parametrize("shape,mult", [((3, 64, 48), 1.0), ((1, 37, 53), 2.5), ((2, 2, 2), 1.0),
                                        ((4, 128, 128), 0.5)])
def test_given_deviates_bit_exact(shape, mult):
    rng = np.random.default_rng(sum(shape))
    depth = rng.uniform(-0.5, 12.0, shape).astype(np.float32)
    depth.reshape(-1)[::97] = np.nan
    normals = rng.standard_normal(shape + (3,)).astype(np.float32) * 2.0
    m = RedwoodDepthNoiseModel(load_model(MODEL), noise_multiplier=mult, seed=1)
    d, n = hip.to_device(depth), hip.to_device(normals)
    out = hip.DeviceBuffer(depth.nbytes)
    m.simulate_device_normals(d.value, shape[0], shape[1], shape[2], n.value, out.value)
    exp = redwood_noise(depth, load_model(MODEL), mult, normals)
    np.testing.assert_array_equal(_host(out, shape), exp)
    assert m.calls == 0  # caller deviates do not advance the generator


def test_generator_blocks_bit_exact():
    seed = 0x0123456789ABCDEF
    m = RedwoodDepthNoiseModel(load_model(MODEL), seed=seed)
    out = hip.DeviceBuffer(64 * 16)
    for call, first in ((0, 0), (7, 1000), (1 << 33, (1 << 32) - 5)):
        m.random_bits_device(call, first, 64, out.value)
        got = _host(out, (64, 4), np.uint32)
        for k in (0, 1, 31, 63):
            g = first + k
            exp = philox4x32([g & 0xFFFFFFFF, g >> 32, call & 0xFFFFFFFF, call >> 32],
                             [seed & 0xFFFFFFFF, seed >> 32])
            np.testing.assert_array_equal(got[k], exp, err_msg=f"call {call} pixel {g}")


def test_zero_noise_is_exact_and_host_equals_device():
    depth = np.linspace(0, 20, num=256 * 256, dtype=np.float32).reshape(256, 256)
    m0 = RedwoodDepthNoiseModel(load_model(MODEL), noise_multiplier=0.0)
    exp = redwood_noise(depth, load_model(MODEL), 0.0, np.zeros((256, 256, 3), np.float32))
    np.testing.assert_array_equal(m0.simulate(depth), exp)
    # the same call index gives the same deviates on the host and device paths
    a = RedwoodDepthNoiseModel(load_model(MODEL), noise_multiplier=1.0, seed=9)
    b = RedwoodDepthNoiseModel(load_model(MODEL), noise_multiplier=1.0, seed=9)
    host = a.simulate(depth)
    d = hip.to_device(depth)
    out = hip.DeviceBuffer(depth.nbytes)
    b.simulate_device(d.value, 1, 256, 256, out.value)
    np.testing.assert_array_equal(_host(out, depth.shape), host)
    assert a.calls == 1 and b.calls == 1
    # the next call draws new deviates; a batch equals its images one by one
    assert not np.array_equal(a.simulate(depth), host)
    pair = np.ascontiguousarray(np.stack([depth, depth[::-1]]))
    c = RedwoodDepthNoiseModel(load_model(MODEL), seed=9)
    dp = hip.to_device(pair)
    op = hip.DeviceBuffer(pair.nbytes)
    c.simulate_device(dp.value, 2, 256, 256, op.value)
    np.testing.assert_array_equal(_host(op, pair.shape)[0], host)


@pytest.mark.parametrize("noise_multiplier,tolerance", [(0.0, 1e-5), (1.0, 1e-2)])
def test_mean_of_simulations_matches_cpu(noise_multiplier, tolerance):
    """tests/test_noise_models.py:24-48: the mean of 100 simulations of a
    0-20 m ramp on the GPU against the mean of 100 CPU simulations (here the
    oracle with numpy's normal deviates)."""
    depth = np.linspace(0, 20, num=256 * 256, dtype=np.float32).reshape(256, 256)
    model = load_model(MODEL)
    m = RedwoodDepthNoiseModel(model, noise_multiplier=noise_multiplier, seed=3)
    rng = np.random.default_rng(0)
    gpu = np.stack([m.simulate(depth) for _ in range(100)])
    cpu = np.stack([redwood_noise(depth, model, noise_multiplier,
                                  rng.standard_normal((256, 256, 3)).astype(np.float32))
                    for _ in range(100)])
    assert np.abs(gpu.mean(0) - cpu.mean(0)).mean() <= tolerance
    # the spread too: the generator's deviates are N(0, 1) like numpy's
    assert np.abs(gpu.std(0) - cpu.std(0)).mean() <= 2 * tolerance


def test_errors():
    with pytest.raises(ValueError):
        RedwoodDepthNoiseModel(np.zeros(10, np.float32))
    m = RedwoodDepthNoiseModel(load_model(MODEL))
    with pytest.raises(N.GbrError) as e:
        m.simulate(np.zeros((1, 8), np.float32))
    assert e.value.status == N.ERR_ARGUMENT
    m.close()


def test_on_the_renderers_linear_depth_plane():
    """The batch renderer's linear-depth plane (one top-down image per env,
    consecutive in device memory) fed straight to the kernel, against the
    oracle's noise over the same observations."""
    from habitat_sim_amd import RendererConfiguration, RendererStandalone, scenes
    spec = scenes.small_rooms(n_envs=4, tile=64)
    gx, gy = spec.grid
    r = RendererStandalone(RendererConfiguration().set_tile_size_count(spec.tile, (gx, gy)),
                           top_down=True)
    scenes.populate(r, spec)
    r.draw()
    r.synchronize()
    n_scenes = gx * gy
    lin = r.plane_device_pointer(N.PLANE_LINEAR_DEPTH)
    rng = np.random.default_rng(8)
    normals = rng.standard_normal((n_scenes, 64, 64, 3)).astype(np.float32)
    nd = hip.to_device(normals)
    out = hip.DeviceBuffer(n_scenes * 64 * 64 * 4)
    m = RedwoodDepthNoiseModel(load_model(MODEL), noise_multiplier=1.0)
    m.simulate_device_normals(lin, n_scenes, 64, 64, nd.value, out.value)
    planes = np.empty((n_scenes, 64, 64), np.float32)
    hip.memcpy_d2h(planes, lin)
    exp = redwood_noise(planes, load_model(MODEL), 1.0, normals)
    np.testing.assert_array_equal(_host(out, planes.shape), exp)
    assert (exp > 0).mean() > 0.5  # the rooms are within 10 m
    r.close()
