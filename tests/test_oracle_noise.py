"""The oracle's Redwood depth noise model (oracle/redwood_ref.c) pinned:
Philox4x32-10 against the published known-answer vectors (Random123,
Salmon et al. SC'11), the noise formula against an independent float64
restatement of the reference's numba model
(redwood_depth_noise_model.py:31-98) with the same deviates, and closed
forms with zero deviates (downsampling, the 10 m cut-off, quantisation to
1/8 disparity, the table's bilinear depth interpolation and its indexing)."""
import os

import numpy as np
import pytest

from oracle.oracle import philox4x32, redwood_noise

MODEL = os.path.join(os.path.dirname(__file__), "golden", "noise", "redwood-depth-dist-model.npy")


def _model():
    return np.load(MODEL, allow_pickle=False).astype(np.float32).reshape(80, 80, 5)


@pytest.mark.parametrize("ctr,key,expected", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])])
def test_philox_known_answers(ctr, key, expected):
    np.testing.assert_array_equal(philox4x32(ctr, key), np.array(expected, np.uint32))


def _numba_model_f64(depth, model, mult, rn):
    """redwood_depth_noise_model.py:_simulate/_undistort in float64, vectorised."""
    H, W = depth.shape
    ymax, xmax = H - 1.0, W - 1.0
    j, i = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    y = (np.clip(j + rn[..., 0] * 0.25 * mult, 0.0, ymax) + 0.5).astype(np.int64)
    x = (np.clip(i + rn[..., 1] * 0.25 * mult, 0.0, xmax) + 0.5).astype(np.int64)
    d = depth[y - y % 2, x - x % 2].astype(np.float64)
    ux = (x / xmax * 639.0 + 0.5).astype(np.int64) // 8
    uy = (y / ymax * 479.0 + 0.5).astype(np.int64) // 6
    i2 = ((d + 1) / 2).astype(np.int64)
    i1 = i2 - 1
    a = (d - (i1 * 2.0 + 1.0)) / 2.0
    f = (1.0 - a) * model[uy, ux, np.clip(i1, 0, 4)] + a * model[uy, ux, np.minimum(i2, 4)]
    und = np.where(f < 1e-5, 0.0, d / np.where(f < 1e-5, 1.0, f))
    den = np.round((35.130 / np.where(und == 0, 1.0, und) + rn[..., 2] * 0.027778 * mult) * 8.0)
    out = np.where(den <= 1e-5, 0.0, 35.130 * 8.0 / np.where(den == 0, 1.0, den))
    out = np.where((d >= 10.0) | (und == 0.0), 0.0, out)
    return out


@pytest.mark.parametrize("mult", [0.0, 1.0, 3.0])
def test_matches_float64_restatement(mult):
    rng = np.random.default_rng(4)
    H, W = 96, 128
    depth = rng.uniform(0.2, 12.0, (H, W)).astype(np.float32)
    rn = rng.standard_normal((H, W, 3)).astype(np.float32)
    ours = redwood_noise(depth, _model(), mult, rn)
    ref = _numba_model_f64(depth, _model(), mult, rn)
    # float32 vs float64 arithmetic: identical but for quantisation steps a
    # rounding apart (and the round-half cases numpy rounds to even)
    close = np.abs(ours - ref) <= 1e-5 * np.maximum(np.abs(ref), 1.0)
    assert close.mean() > 0.998, close.mean()
    assert (ours[depth[::2, ::2].repeat(2, 0).repeat(2, 1) >= 10.0] == 0).all() if mult == 0 else True


def test_closed_forms_with_zero_deviates():
    H, W = 48, 64
    rng = np.random.default_rng(2)
    depth = rng.uniform(0.5, 9.5, (H, W)).astype(np.float32)
    depth[4, 6] = 12.0   # beyond 10 m (read by its whole 2x2 block)
    depth[10, 20] = 0.0  # no return
    ones = np.ones((80, 80, 5), np.float32)
    out = redwood_noise(depth, ones, 1.0, np.zeros((H, W, 3), np.float32))
    # every pixel reads the top-left pixel of its 2x2 block
    blk = depth[::2, ::2].repeat(2, 0).repeat(2, 1)
    with np.errstate(divide="ignore"):
        q = np.round(35.130 / blk.astype(np.float64) * 8.0)
    exp = (np.float32(35.130) * 8.0 / q).astype(np.float32)
    exp[(blk >= 10.0) | (blk == 0.0)] = 0.0
    np.testing.assert_array_equal(out, exp)
    assert (out[4:6, 6:8] == 0).all() and (out[10:12, 20:22] == 0).all()


def test_table_interpolation_and_indexing():
    """A table whose entries encode their own indices: at zero deviates the
    undistorted depth is z / ((1-a) M[r][c][i1] + a M[r][c][i2]) with the
    table cell of the pixel's 640x480-scaled position."""
    H, W = 60, 80
    r, c, k = np.meshgrid(np.arange(80), np.arange(80), np.arange(5), indexing="ij")
    model = (0.5 + 0.005 * r + 0.003 * c + 0.05 * k).astype(np.float32)
    depth = np.full((H, W), 3.3, np.float32)
    out = redwood_noise(depth, model, 0.0, np.zeros((H, W, 3), np.float32))
    j, i = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    cx = ((i / (W - 1.0)) * 639.0 + 0.5).astype(int) // 8
    cy = ((j / (H - 1.0)) * 479.0 + 0.5).astype(int) // 6
    z = 3.3
    i2 = int((z + 1) / 2)  # 2
    a = (z - (2 * (i2 - 1) + 1)) / 2.0
    f = (1 - a) * model[cy, cx, i2 - 1].astype(np.float64) + a * model[cy, cx, i2].astype(np.float64)
    exp = 35.130 * 8.0 / np.round(35.130 / (z / f) * 8.0)
    np.testing.assert_allclose(out, exp, rtol=1e-6)
    assert len(np.unique(out)) > 10  # the cells differ


def test_model_fixture_shape():
    m = _model()
    assert m.shape == (80, 80, 5) and np.isfinite(m).all() and m.max() < 2.0
