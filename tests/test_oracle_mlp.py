"""Pins the CPU oracle of the chunk-visibility MLP (oracle/mlp_ref.py):
bfloat16 rounding known answers, the float64 gradient against central finite
differences, and the rounded model against the unrounded one."""
import numpy as np
import pytest

from oracle import mlp_ref as M


def test_bf16_round_to_nearest_even():
    assert M.bf16(1.0) == 1.0
    assert M.bf16(1 + 2.0 ** -8) == 1.0                  # tie -> even (down)
    assert M.bf16(1 + 3 * 2.0 ** -8) == 1 + 2.0 ** -6    # tie -> even (up)
    assert M.bf16(1 + 2.0 ** -8 + 2.0 ** -20) == 1 + 2.0 ** -7  # above the tie
    assert M.bf16(-1.5) == -1.5
    x = np.random.default_rng(0).normal(size=1000)
    r = M.bf16(x)
    # 8 significant bits: rounding error at most half an ulp, 2^-8 relative
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)
    assert np.all(M.bf16(r) == r)  # idempotent


def _data(n=64, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, M.IN))
    y = (rng.uniform(size=n) < 0.3).astype(np.uint8)
    return X, y


@pytest.mark.parametrize("pos_weight", [1.0, 3.0])
def test_gradient_matches_finite_differences(pos_weight):
    flat = M.he_uniform(3)
    X, y = _data()
    _, g = M.loss_grad(flat, X, y, pos_weight, rounding=False)
    rng = np.random.default_rng(1)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in M.SIZES])
    idx = np.concatenate([rng.integers(offs[i], offs[i + 1], 8) for i in range(len(M.SIZES))])
    for i in idx:
        h = 1e-6 * max(1.0, abs(flat[i]))
        a, b = flat.copy(), flat.copy()
        a[i] += h
        b[i] -= h
        fd = (M.loss_only(a, X, y, pos_weight) - M.loss_only(b, X, y, pos_weight)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_loss_closed_form_at_zero_weights():
    """All weights zero and b3 = c: every logit is c, so the loss is the
    weighted BCE of a constant and only b3 has a gradient, w*(sigmoid(c)-y)."""
    flat = np.zeros(M.PARAMS)
    c = 0.7
    flat[-1] = c
    X, y = _data(32, 5)
    loss, g = M.loss_grad(flat, X, y, 2.0, rounding=False)
    w = np.where(y > 0, 2.0, 1.0)
    ref = np.mean(w * (np.log1p(np.exp(-c)) + c - c * y))
    assert abs(loss - ref) < 1e-12
    sig = 1 / (1 + np.exp(-c))
    assert abs(g[-1] - np.mean(w * (sig - y))) < 1e-12
    assert np.all(g[:-1] == 0)


def test_rounded_model_tracks_float64_model():
    flat = M.he_uniform(7)
    X, y = _data(128, 2)
    z0, _ = M.forward(flat, X, rounding=False)
    z1, _ = M.forward(flat, X, rounding=True)
    assert np.max(np.abs(z1 - z0)) < 0.05 * (np.max(np.abs(z0)) + 1e-3)
    _, g0 = M.loss_grad(flat, X, y, rounding=False)
    _, g1 = M.loss_grad(flat, X, y, rounding=True)
    assert np.linalg.norm(g1 - g0) < 0.05 * np.linalg.norm(g0)
