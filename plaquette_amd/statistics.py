"""FP64 statistics of measurement series (statistics.rs:83-317).

The one FP64 step of the path: the device records integer sums, and these
functions turn the series into estimates.

- ``reduce``: mean, standard error, integrated autocorrelation time and
  effective sample count. tau_int uses Sokal's self-consistent window, c = 5
  (statistics.rs:83-137). The autocovariance comes from an FFT for long
  series, and from direct sums otherwise.
- ``specific_heat``, ``susceptibility``, ``binder``: blocked jackknife
  estimates (statistics.rs:152-284). The block length is ceil(2 tau_int) of
  the slowest raw series.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_FFT_MIN = 4096


@dataclass
class Estimate:
    """statistics.rs:22-41."""
    mean: float
    stderr: float
    tau_int: float
    n_eff: float


@dataclass
class Derived:
    """statistics.rs:43-58: a jackknifed value, its error and the block count."""
    value: float
    stderr: float
    n_blocks: int


def _autocov(d: np.ndarray) -> np.ndarray:
    """c[t] = sum_i d[i] d[i + t] / n for t = 0 .. n - 1."""
    n = d.size
    if n >= _FFT_MIN:
        size = 1 << int(math.ceil(math.log2(2 * n)))
        f = np.fft.rfft(d, size)
        return np.fft.irfft(f * np.conj(f), size)[:n] / n
    return np.array([np.dot(d[: n - t], d[t:]) for t in range(n)]) / n


def reduce(series) -> Estimate:
    """statistics.rs:83-137."""
    x = np.asarray(series, dtype=np.float64).ravel()
    n = x.size
    if n < 2:
        return Estimate(float(x[0]) if n == 1 else 0.0, 0.0, 0.5, float(n))
    mean = float(x.mean())
    d = x - mean
    c0 = float(np.dot(d, d)) / n
    if c0 == 0.0:
        return Estimate(mean, 0.0, 0.5, float(n))
    rho = _autocov(d)[1:] / c0
    taus = 0.5 + np.cumsum(rho)                    # tau after lag t = 1 .. n - 1
    t = np.arange(1, n, dtype=np.float64)
    stop = np.nonzero(t >= 5.0 * taus)[0]
    tau = float(taus[stop[0]] if stop.size else taus[-1])
    tau = max(tau, 0.5)
    return Estimate(mean, math.sqrt(2.0 * tau * c0 / n), tau, n / (2.0 * tau))


def _jackknife(features, block_len: int, estimator) -> Derived:
    """statistics.rs:152-211: leave-one-block-out estimates."""
    feats = [np.asarray(f, dtype=np.float64) for f in features]
    n = feats[0].size
    value = estimator([float(f.mean()) for f in feats])
    block_len = max(int(block_len), 1)
    nb = n // block_len
    if nb < 2:
        return Derived(value, math.nan, nb)
    used = nb * block_len
    partial = [f[:used].reshape(nb, block_len).sum(axis=1) for f in feats]
    totals = [float(p.sum()) for p in partial]
    leave_n = float(used - block_len)
    thetas = np.array([estimator([(totals[j] - partial[j][b]) / leave_n for j in range(len(feats))])
                       for b in range(nb)])
    spread = float(np.sum((thetas - thetas.mean()) ** 2))
    return Derived(value, math.sqrt((nb - 1.0) / nb * spread), nb)


def _block_len(*series) -> int:
    """statistics.rs:216-219: ceil(2 tau_int) of the slowest series."""
    tau = max([0.5] + [reduce(s).tau_int for s in series])
    return max(int(math.ceil(2.0 * tau)), 1)


def specific_heat(energies, beta: float, n_sites: float) -> Derived:
    """statistics.rs:232-244: C = beta^2 Var(E) / N."""
    e = np.asarray(energies, dtype=np.float64)
    f0 = e - e.mean()
    bl = _block_len(e, e * e)
    return _jackknife([f0, f0 * f0], bl, lambda m: beta * beta * max(m[1] - m[0] * m[0], 0.0) / n_sites)


def susceptibility(magnetisations, beta: float, n_sites: float) -> Derived:
    """statistics.rs:248-264: chi = beta N Var(|m|), m = M / N."""
    a = np.abs(np.asarray(magnetisations, dtype=np.float64) / n_sites)
    f0 = a - a.mean()
    bl = _block_len(a, a * a)
    return _jackknife([f0, f0 * f0], bl, lambda m: beta * n_sites * max(m[1] - m[0] * m[0], 0.0))


def binder(magnetisations) -> Derived:
    """statistics.rs:277-284: U4 = 1 - <m^4> / (3 <m^2>^2)."""
    m = np.asarray(magnetisations, dtype=np.float64)
    m2 = m * m
    m4 = m2 * m2
    bl = _block_len(m2, m4)
    return _jackknife([m2, m4], bl, lambda v: 1.0 - v[1] / (3.0 * v[0] * v[0]))
