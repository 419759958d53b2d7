"""FP64 numpy restatement of the warm-start network -- TEST INFRASTRUCTURE ONLY.

Same architecture, parameter layout, normalisation and loss as the HIP
implementation (torchmorph_amd/csrc/tm_warmstart.hip), with a hand-written
backward.  ``bf16=True`` rounds (round-to-nearest-even) the same operands the
GPU stores in bfloat16 -- network inputs, the GEMM weights W0, W1, W2a, W3,
the activations and the GEMM-side upstream gradients -- so the two agree to
accumulation-order effects.  The backward is pinned by central differences in
tests/test_oracle.py.
"""
from __future__ import annotations

import numpy as np

K_IN, H, G = 4, 128, 256
LAYOUT = [("W0", (K_IN, H)), ("b0", (H,)), ("W1", (H, H)), ("b1", (H,)),
          ("W2a", (H, G)), ("W2x", (H, G)), ("W2y", (H, G)), ("w2l", (G,)), ("b2", (G,)),
          ("W3", (G, G)), ("b3", (G,)), ("w4", (G,)), ("b4", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in LAYOUT)


def unpack(flat):
    out, o = {}, 0
    for name, shape in LAYOUT:
        k = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + k], np.float64).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).reshape(-1) for name, _ in LAYOUT])


def bf16(a):
    """Round float values to bfloat16 (RNE) and back to float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _stats(x, y, eps):
    c = y.mean(0)
    s2 = ((x - c) ** 2).sum(1).mean() + ((y - c) ** 2).sum(1).mean()
    return c, s2, np.log(eps / s2)


def forward(params, pairs, round_bf16=False):
    """pairs: list of (x (n,3), y (m,3), eps).  Returns (outputs per pair in
    units of s2, cache)."""
    R = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    P = unpack(params)
    W0, W1, W2a, W3 = R(P["W0"]), R(P["W1"]), R(P["W2a"]), R(P["W3"])
    relu = lambda a: np.maximum(a, 0.0)
    outs, cache = [], []
    for x, y, eps in pairs:
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        c, s2, r = _stats(x, y, eps)
        sc = 1.0 / np.sqrt(s2)
        Ux = R(np.concatenate([(x - c) * sc, np.zeros((len(x), 1))], 1))
        Uy = R(np.concatenate([(y - c) * sc, np.ones((len(y), 1))], 1))
        H1x = R(relu(Ux @ W0 + P["b0"]))
        H1y = R(relu(Uy @ W0 + P["b0"]))
        H2x = R(relu(H1x @ W1 + P["b1"]))
        H2y = R(relu(H1y @ W1 + P["b1"]))
        ex, ey = H2x.mean(0), H2y.mean(0)
        T = ex @ P["W2x"] + ey @ P["W2y"] + r * P["w2l"] + P["b2"]
        A1 = R(relu(H2x @ W2a + T))
        A2 = R(relu(A1 @ W3 + P["b3"]))
        out = A2 @ P["w4"] + P["b4"][0]
        outs.append(out)
        cache.append(dict(Ux=Ux, Uy=Uy, H1x=H1x, H1y=H1y, H2x=H2x, H2y=H2y, ex=ex, ey=ey, r=r,
                          A1=A1, A2=A2, s2=s2))
    return outs, cache


def predict(params, pairs, round_bf16=False):
    outs, cache = forward(params, pairs, round_bf16)
    return [o * c["s2"] for o, c in zip(outs, cache)]


def loss_and_grad(params, pairs, targets, round_bf16=False, alpha=0.0):
    """Centred MSE (mean over pairs of (s2/eps)^alpha times the per-pair mean
    of squared, mean-removed residuals in units of s2) and its gradient."""
    R = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    P = unpack(params)
    W1, W2a, W3 = R(P["W1"]), R(P["W2a"]), R(P["W3"])
    outs, cache = forward(params, pairs, round_bf16)
    npairs = len(pairs)
    g = {name: np.zeros(shape) for name, shape in LAYOUT}
    loss = 0.0
    for out, cc, t, pr in zip(outs, cache, targets, pairs):
        res = out - np.asarray(t, np.float64) / cc["s2"]
        res = res - res.mean()
        n = len(out)
        wp = (cc["s2"] / pr[2]) ** alpha
        loss += wp * (res ** 2).sum() / n / npairs
        dout = 2.0 * wp * res / (n * npairs)
        g["w4"] += cc["A2"].T @ dout
        g["b4"] += dout.sum()
        dA2 = np.outer(dout, P["w4"])
        dZ2 = dA2 * (cc["A2"] > 0)
        g["b3"] += dZ2.sum(0)
        dZ2b = R(dZ2)
        g["W3"] += cc["A1"].T @ dZ2b
        dA1 = dZ2b @ W3.T
        dZ1 = dA1 * (cc["A1"] > 0)
        dZ1b = R(dZ1)
        g["W2a"] += cc["H2x"].T @ dZ1b
        dH2x = dZ1b @ W2a.T
        dT = dZ1.sum(0)
        g["W2x"] += np.outer(cc["ex"], dT)
        g["W2y"] += np.outer(cc["ey"], dT)
        g["w2l"] += cc["r"] * dT
        g["b2"] += dT
        dex = P["W2x"] @ dT
        dey = P["W2y"] @ dT
        dH2x = dH2x + dex / len(cc["H2x"])
        dH2y = np.tile(dey / len(cc["H2y"]), (len(cc["H2y"]), 1))
        dZx = dH2x * (cc["H2x"] > 0)
        dZy = dH2y * (cc["H2y"] > 0)
        g["b1"] += dZx.sum(0) + dZy.sum(0)
        dZxb, dZyb = R(dZx), R(dZy)
        g["W1"] += cc["H1x"].T @ dZxb + cc["H1y"].T @ dZyb
        dH1x = dZxb @ W1.T
        dH1y = dZyb @ W1.T
        dZx = dH1x * (cc["H1x"] > 0)
        dZy = dH1y * (cc["H1y"] > 0)
        g["b0"] += dZx.sum(0) + dZy.sum(0)
        g["W0"] += cc["Ux"].T @ R(dZx) + cc["Uy"].T @ R(dZy)
    return loss, pack(g)


def init_params(seed=0):
    """He-normal weights, zero biases (same rule as the package's init)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in LAYOUT:
        if len(shape) == 2:
            d[name] = rng.normal(0, np.sqrt(2.0 / shape[0]), size=shape)
        elif name in ("w2l", "w4"):
            d[name] = rng.normal(0, np.sqrt(1.0 / shape[0]), size=shape)
        else:
            d[name] = np.zeros(shape)
    return pack(d)
