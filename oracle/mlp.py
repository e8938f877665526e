"""CPU oracle of the radiance-cache MLP — TEST INFRASTRUCTURE ONLY.

Only tests/ (and __graft_entry__.smoke()) may use this module, as the
checker of the GPU kernels in hypertrace_amd/csrc/cache.hip. The reference
has no learned part, so this is a restatement of the method itself:

  encode   12 features x 6 bands -> 144 inputs, index 12 f + 2 b + c:
           c = 0: sin(2^b pi x_f), c = 1: cos(2^b pi x_f)
  forward  144 -> 64 -> 64 -> 64 -> 64 -> 64 (ReLU) -> 3 (linear)
  loss     mean over samples x channels of (p - y)^2 / (sg(p)^2 + 0.01)
  backward the analytic gradient of that loss (denominator held fixed)
  adam     beta1 0.9, beta2 0.99, eps 1e-8, bias-corrected

`emulate=True` rounds what the GPU rounds to BF16 (round to nearest even):
the encoded inputs, the weights, each hidden activation and each
back-propagated gradient; accumulation is FP64 (the GPU accumulates in FP32,
so agreement is to a tolerance). `emulate=False` is exact FP64, pinned by
finite differences (tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import numpy as np

LOSS_EPS = 0.01
ADAM = (0.9, 0.99, 1e-8)
NAMES = ["W0", "b0", "W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4", "W5", "b5"]


def bf16(x) -> np.ndarray:
    """Round float32 values to the nearest BF16 (ties to even), as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def encode(feat: np.ndarray, emulate: bool = True) -> np.ndarray:
    feat = np.asarray(feat, np.float64)
    n = feat.shape[0]
    out = np.zeros((n, 144))
    for f in range(12):
        for b in range(6):
            a = (2.0 ** b) * np.pi * feat[:, f]
            out[:, 12 * f + 2 * b] = np.sin(a)
            out[:, 12 * f + 2 * b + 1] = np.cos(a)
    return bf16(out) if emulate else out


def forward(p: dict, feat: np.ndarray, emulate: bool = True) -> dict:
    r = (lambda v: bf16(v)) if emulate else (lambda v: np.asarray(v, np.float64))
    x = encode(feat, emulate)
    acts = [x]
    h = x
    for l in range(5):
        y = h @ r(p[f"W{l}"]).T + np.asarray(p[f"b{l}"], np.float64)
        h = r(np.maximum(y, 0.0))
        acts.append(h)
    out = h @ r(p["W5"]).T + np.asarray(p["b5"], np.float64)
    return {"acts": acts, "out": out}


def loss(out: np.ndarray, target: np.ndarray, den=None) -> float:
    den = out * out + LOSS_EPS if den is None else den
    return float(np.mean((out - target) ** 2 / den))


def backward(p: dict, fwd: dict, target: np.ndarray, emulate: bool = True) -> dict:
    r = (lambda v: bf16(v)) if emulate else (lambda v: np.asarray(v, np.float64))
    out = fwd["out"]
    n = out.shape[0]
    if emulate:
        o32 = out.astype(np.float32)
        d = (2.0 * (o32 - target.astype(np.float32)) / ((o32 * o32 + np.float32(LOSS_EPS)) * np.float32(n * 3)))
        d = bf16(d)
    else:
        d = 2.0 * (out - target) / ((out * out + LOSS_EPS) * n * 3)
    g = {}
    acts = fwd["acts"]
    for l in range(5, -1, -1):
        a = acts[l]
        g[f"W{l}"] = d.T @ a
        g[f"b{l}"] = d.sum(axis=0)
        if l > 0:
            da = d @ r(p[f"W{l}"])
            d = r(da * (acts[l] > 0.0))
    return g


def adam(p: dict, m: dict, v: dict, g: dict, step: int, lr: float) -> None:
    b1, b2, eps = ADAM
    c1 = 1.0 / (1.0 - b1 ** step)
    c2 = 1.0 / (1.0 - b2 ** step)
    for k in NAMES:
        m[k] = b1 * m[k] + (1.0 - b1) * g[k]
        v[k] = b2 * v[k] + (1.0 - b2) * g[k] * g[k]
        p[k] = p[k] - lr * (m[k] * c1) / (np.sqrt(v[k] * c2) + eps)
