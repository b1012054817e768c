"""CPU oracle of the warm-start MLP (24 -> 64 -> 64 -> 4, ReLU) — TEST
INFRASTRUCTURE ONLY.  float64 NumPy forward/backward; ``bf16=True`` rounds the
operands the GPU feeds to its BF16 matrix instructions (layers 1-2 and the
backward products), so GPU and oracle see the same inputs."""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 24, 64, 4


def bf16_round(x):
    """round-to-nearest-even to bfloat16, returned as float64"""
    x32 = np.ascontiguousarray(x, np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    u = ((u + (((u >> 16) & 1) + 0x7FFF)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(x):
    x32 = np.ascontiguousarray(x, np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    return (((u + (((u >> 16) & 1) + 0x7FFF)) >> 16) & 0xFFFF).astype(np.uint16)


def init_params(rng, scale=1.0):
    return {
        "W1": rng.normal(0, scale * np.sqrt(2.0 / NIN), (NH, NIN)),
        "b1": rng.normal(0, 0.05, NH),
        "W2": rng.normal(0, scale * np.sqrt(2.0 / NH), (NH, NH)),
        "b2": rng.normal(0, 0.05, NH),
        "W3": rng.normal(0, scale * np.sqrt(1.0 / NH), (NOUT, NH)),
        "b3": rng.normal(0, 0.05, NOUT),
    }


def forward(p, X, bf16=True):
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    X = np.asarray(X, np.float64)
    H1 = np.maximum(r(X) @ r(p["W1"]).T + p["b1"], 0.0)
    H2 = np.maximum(r(H1) @ r(p["W2"]).T + p["b2"], 0.0)
    Y = H2 @ np.asarray(p["W3"], np.float64).T + p["b3"]
    return Y, H1, H2


def backward(p, X, H1, H2, dY, bf16=True):
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    g = {}
    g["W3"] = dY.T @ H2
    g["b3"] = dY.sum(0)
    dH2 = (dY @ np.asarray(p["W3"], np.float64)) * (H2 > 0)
    g["W2"] = r(dH2).T @ r(H1)
    g["b2"] = dH2.sum(0)
    dH1 = (r(dH2) @ r(p["W2"])) * (H1 > 0)
    g["W1"] = r(dH1).T @ r(X)
    g["b1"] = dH1.sum(0)
    return g


def flatten(g):
    return np.concatenate([np.ravel(g[k]) for k in ("W1", "b1", "W2", "b2", "W3", "b3")])
