"""CPU oracle of the learned score floor's MLP -- TEST INFRASTRUCTURE ONLY.

A float64 numpy restatement of the network the HIP kernels (hybridsearchsystem_amd/csrc/frk_mlp.hip)
train and run: x -> relu(W1 x + b1) -> relu(W2 h1 + b2) -> w3 . h2 + b3, with a pinball (quantile)
loss. With ``round_bf16=True`` every matrix product rounds its two operands to BF16 (nearest even)
exactly where the kernels do, and values the kernels store in FP32 are rounded to FP32, so the GPU
results should match to FP32 accumulation-order error; the backward pass treats the roundings as
identity (the straight-through gradient the kernels compute). With ``round_bf16=False`` it is the
exact gradient of the float64 network, which the tests pin with central finite differences.

There is no reference counterpart: the reference has no learned component.
"""
import numpy as np


def bf16(x):
    """Round to BF16 (nearest even), returned as float32 values."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000  # inf keeps its bits, NaN stays NaN
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    quiet = np.where((u & 0xFFFF) != 0, np.uint64(0x40), np.uint64(0))
    r = np.where(special, ((u >> np.uint64(16)) | quiet) << np.uint64(16), r)
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def _round(x, on):
    return bf16(x).astype(np.float64) if on else np.asarray(x, dtype=np.float64)


def _f32(x, on):
    return np.asarray(x, dtype=np.float32).astype(np.float64) if on else x


def init_params(in_dim, hidden, seed=0):
    """He-uniform init (float32), the same rule the product trainer uses for its starting point."""
    rng = np.random.default_rng(seed)

    def lin(fan_out, fan_in):
        lim = np.sqrt(6.0 / fan_in)
        return rng.uniform(-lim, lim, size=(fan_out, fan_in)).astype(np.float32)

    return {"w1": lin(hidden, in_dim), "b1": np.zeros(hidden, np.float32), "w2": lin(hidden, hidden),
            "b2": np.zeros(hidden, np.float32), "w3": lin(1, hidden).reshape(hidden),
            "b3": np.zeros(1, np.float32)}


def forward(p, x, round_bf16=True):
    r = round_bf16
    z1 = _f32(_round(x, r) @ _round(p["w1"], r).T + np.asarray(p["b1"], np.float64), r)
    h1 = np.maximum(z1, 0.0)
    z2 = _f32(_round(h1, r) @ _round(p["w2"], r).T + np.asarray(p["b2"], np.float64), r)
    h2 = np.maximum(z2, 0.0)
    y = _f32(_round(h2, r) @ _round(p["w3"], r) + float(p["b3"][0]), r)
    return {"x": np.asarray(x, np.float64), "z1": z1, "h1": h1, "z2": z2, "h2": h2, "y": y}


def pinball(y, label, tau):
    """Mean pinball loss of r = label - y and its gradient in y."""
    y = np.asarray(y, np.float64)
    r = np.asarray(label, np.float64) - y
    loss = np.where(r >= 0, tau * r, (tau - 1.0) * r)
    dy = np.where(r > 0, -tau, 1.0 - tau) / len(y)
    return float(loss.mean()), dy


def backward(p, c, dy, round_bf16=True):
    r = round_bf16
    dy = np.asarray(dy, np.float64)
    w3r = _round(p["w3"], r)
    dz2 = _f32(np.where(c["z2"] > 0, dy[:, None] * w3r[None, :], 0.0), r)
    g = {"w3": _f32(dy @ _round(c["h2"], r), r), "b3": np.array([dy.sum()])}
    g["w2"] = _f32(_round(dz2, r).T @ _round(c["h1"], r), r)
    g["b2"] = _f32(dz2.sum(axis=0), r)
    dh1 = _round(dz2, r) @ _round(p["w2"], r)
    dz1 = _f32(np.where(c["z1"] > 0, dh1, 0.0), r)
    g["w1"] = _f32(_round(dz1, r).T @ _round(c["x"], r), r)
    g["b1"] = _f32(dz1.sum(axis=0), r)
    return g


def loss_and_grads(p, x, label, tau, round_bf16=True):
    c = forward(p, x, round_bf16)
    loss, dy = pinball(c["y"], label, tau)
    return loss, backward(p, c, dy, round_bf16), c
