"""Exact partition function of the 2D Ising model on an m x n torus.

Kaufman's solution in the form of Ferdinand & Fisher, Phys. Rev. 185, 832
(1969), eqs. (2.1)-(2.5), with J = 1, K = beta:

    Z = 1/2 (2 sinh 2K)^{mn/2} (Z1 + Z2 + Z3 + Z4)
    Z1 = prod_{r<n} 2 cosh(m g_{2r+1} / 2)     Z2 = prod_{r<n} 2 sinh(m g_{2r+1} / 2)
    Z3 = prod_{r<n} 2 cosh(m g_{2r} / 2)       Z4 = prod_{r<n} 2 sinh(m g_{2r} / 2)
    cosh g_l = cosh 2K coth 2K - cos(pi l / n)  (l >= 1),   g_0 = 2K + ln tanh K

g_0 is negative below the critical temperature, which makes Z4 change sign.
Everything is summed in logarithms.  Test infrastructure: it pins the
sampler's observables, and is itself pinned against enumeration.
"""
from __future__ import annotations

import math

import numpy as np


def _gammas(K: float, n: int) -> np.ndarray:
    l = np.arange(2 * n, dtype=np.float64)
    c = math.cosh(2 * K) / math.tanh(2 * K) - np.cos(np.pi * l / n)
    g = np.arccosh(np.maximum(c, 1.0))
    g[0] = 2 * K + math.log(math.tanh(K))
    return g


def _log_prod(terms: np.ndarray):
    """(log |prod|, sign) of a product of real terms."""
    sign = 1.0 if np.count_nonzero(terms < 0) % 2 == 0 else -1.0
    return float(np.sum(np.log(np.abs(terms)))), sign


def log_z(m: int, n: int, beta: float) -> float:
    K = beta
    g = _gammas(K, n)
    odd = g[1::2][:n]
    even = g[0::2][:n]
    parts = [_log_prod(2 * np.cosh(m * odd / 2)), _log_prod(2 * np.sinh(m * odd / 2)),
             _log_prod(2 * np.cosh(m * even / 2)), _log_prod(2 * np.sinh(m * even / 2))]
    top = max(p[0] for p in parts)
    s = sum(sign * math.exp(lp - top) for lp, sign in parts)
    return math.log(0.5) + 0.5 * m * n * math.log(2 * math.sinh(2 * K)) + top + math.log(s)


def energy_per_site(m: int, n: int, beta: float, h: float = 1e-5) -> float:
    """<E>/N = -(d ln Z / d beta) / N by a central difference."""
    return -(log_z(m, n, beta + h) - log_z(m, n, beta - h)) / (2 * h) / (m * n)


def specific_heat(m: int, n: int, beta: float, h: float = 1e-4) -> float:
    """C = beta^2 (d^2 ln Z / d beta^2) / N, the convention of statistics.rs:232-244."""
    d2 = (log_z(m, n, beta + h) - 2 * log_z(m, n, beta) + log_z(m, n, beta - h)) / (h * h)
    return beta * beta * d2 / (m * n)


def enumerate_log_z(m: int, n: int, beta: float):
    """Brute force over all 2^(mn) states: (ln Z, <E>/N)."""
    N = m * n
    states = ((np.arange(1 << N)[:, None] >> np.arange(N)[None, :]) & 1).astype(np.int8)
    s = 1 - 2 * states
    grid = s.reshape(-1, n, m)
    bonds = (grid * np.roll(grid, -1, axis=2)).sum(axis=(1, 2)) + (grid * np.roll(grid, -1, axis=1)).sum(axis=(1, 2))
    e = -bonds.astype(np.float64)
    w = -beta * e
    top = w.max()
    z = np.exp(w - top).sum()
    return top + math.log(z), float((e * np.exp(w - top)).sum() / z / N)
