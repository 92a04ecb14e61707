"""Ring-only logits for checking the block Metropolis-Hastings step.

Test infrastructure.  The block MH step is exact for any logits that depend
on a block's boundary ring only, so its stationary law can be checked with
logits that need no network.  These are the naive mean-field logits of the
8x8 interior given the ring: m = tanh(beta h) iterated to a fixed point,
h = J (the sum of the four neighbours' magnetisations, ring spins fixed) +
the external field, and logit = 2 beta h, the log-odds of state 0 (spin +1)
for a spin in field h.
"""
from __future__ import annotations

import numpy as np


def _patch_index(L0: int, L1: int) -> np.ndarray:
    """[block, 10, 10] lattice indices of each block's padded patch."""
    nb0, nb1 = L0 // 8, L1 // 8
    idx = np.zeros((nb1, nb0, 10, 10), np.int64)
    for by in range(nb1):
        ys = (by * 8 - 1 + np.arange(10)) % L1
        for bx in range(nb0):
            xs = (bx * 8 - 1 + np.arange(10)) % L0
            idx[by, bx] = ys[:, None] * L0 + xs[None, :]
    return idx.reshape(nb0 * nb1, 10, 10)


_CACHE: dict = {}


def mean_field_logits(shape, state, beta: float, j: float = 1.0, h: float = 0.0, iters: int = 30) -> np.ndarray:
    """(n_blocks, 64) float32 logits, block index by * L0/8 + bx, site py * 8 + px."""
    L0, L1 = shape
    if shape not in _CACHE:
        _CACHE[shape] = _patch_index(L0, L1)
    idx = _CACHE[shape]
    s = 1.0 - 2.0 * np.asarray(state, np.float64)[idx]
    s[:, 1:9, 1:9] = 0.0
    s[:, 0, 0] = s[:, 0, 9] = s[:, 9, 0] = s[:, 9, 9] = 0.0
    f = s.copy()
    field = np.zeros((idx.shape[0], 8, 8))
    for _ in range(iters):
        field = j * (f[:, 0:8, 1:9] + f[:, 2:10, 1:9] + f[:, 1:9, 0:8] + f[:, 1:9, 2:10]) + h
        f[:, 1:9, 1:9] = np.tanh(beta * field)
    return (2.0 * beta * field).reshape(-1, 64).astype(np.float32)
