"""TEST INFRASTRUCTURE ONLY - the CPU oracle of the chunk-visibility MLP.

Nothing in the product imports this module; tests use it to check the GPU
forward and backward passes (habitat-sim_amd/csrc/gbr_mlp.hip, gbr_vis.cpp).
The model has no counterpart in the reference (its batch renderer draws
every mesh, src/esp/gfx_batch/Renderer.cpp:1298-1351); it is the learned
proposal of the two-phase chunk cull: 32 -> 256 -> 256 -> 1, ReLU, weighted
binary cross-entropy with logits.

Parameter layout (flat float32, 74497 values):
  W1 [256][32], b1 [256], W2 [256][256], b2 [256], w3 [256], b3

``rounding=True`` rounds to bfloat16 (round to nearest even) exactly where
the GPU does: inputs, weights W1 W2 w3, hidden activations H1 and H2 (after
bias and ReLU), and the back-propagated dH2 and dH1; sums are float64 here
and float32 on the GPU.  ``rounding=False`` is the plain float64 model,
whose gradient the tests pin against central finite differences.
"""
import numpy as np

IN, HID = 32, 256
SIZES = (("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
         ("w3", (HID,)), ("b3", (1,)))
PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(flat):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES])


def bf16(x):
    """Round to bfloat16 (nearest even) through float32; returns float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _r(x, rounding):
    return bf16(x) if rounding else np.asarray(x, np.float64)


def forward(flat, X, rounding=True):
    """logits and the activations the backward pass needs."""
    p = unpack(flat)
    Xb = _r(X, rounding)
    W1, W2, w3 = _r(p["W1"], rounding), _r(p["W2"], rounding), _r(p["w3"], rounding)
    H1 = _r(np.maximum(Xb @ W1.T + p["b1"], 0.0), rounding)
    H2 = _r(np.maximum(H1 @ W2.T + p["b2"], 0.0), rounding)
    z = H2 @ w3 + p["b3"][0]
    return z, dict(Xb=Xb, W2=W2, w3=w3, H1=H1, H2=H2)


def loss_grad(flat, X, y, pos_weight=1.0, rounding=True):
    """Mean weighted BCE-with-logits loss and its gradient (flat, float64).

    loss = mean_r w_r [log(1 + e^-|z_r|) + max(z_r, 0) - z_r y_r],
    w_r = pos_weight if y_r else 1."""
    y = np.asarray(y, np.float64)
    n = len(y)
    z, a = forward(flat, X, rounding)
    w = np.where(y > 0, pos_weight, 1.0)
    loss = np.sum(w * (np.log1p(np.exp(-np.abs(z))) + np.maximum(z, 0.0) - z * y)) / n
    sig = 1.0 / (1.0 + np.exp(-z))
    dl = w * (sig - y) / n
    g = {"b3": np.array([dl.sum()]), "w3": a["H2"].T @ dl}
    dH2 = _r(np.where(a["H2"] > 0, dl[:, None] * a["w3"][None, :], 0.0), rounding)
    g["W2"] = dH2.T @ a["H1"]
    g["b2"] = dH2.sum(0)
    dH1 = _r(np.where(a["H1"] > 0, dH2 @ a["W2"], 0.0), rounding)
    g["W1"] = dH1.T @ a["Xb"]
    g["b1"] = dH1.sum(0)
    return loss, pack(g)


def loss_only(flat, X, y, pos_weight=1.0):
    y = np.asarray(y, np.float64)
    z, _ = forward(flat, X, rounding=False)
    w = np.where(y > 0, pos_weight, 1.0)
    return np.sum(w * (np.log1p(np.exp(-np.abs(z))) + np.maximum(z, 0.0) - z * y)) / len(y)


def he_uniform(seed):
    """Random weights of the same family the library initialises with."""
    rng = np.random.default_rng(seed)
    d = {"W1": rng.uniform(-1, 1, (HID, IN)) * np.sqrt(6.0 / IN),
         "b1": rng.normal(0, 0.1, HID),
         "W2": rng.uniform(-1, 1, (HID, HID)) * np.sqrt(6.0 / HID),
         "b2": rng.normal(0, 0.1, HID),
         "w3": rng.uniform(-1, 1, HID) * np.sqrt(6.0 / HID) * 0.1,
         "b3": np.array([-0.5])}
    return pack(d)
