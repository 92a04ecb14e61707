"""Exact results for small Z2 gauge planes and the replica-exchange check.

Test infrastructure.  The density of states of the signed plaquette sum of a
small periodic plane is found by enumerating every link configuration, so
the equilibrium law of the plaquette sum at any beta is exact.  Replica
exchange between two replicas at beta_k < beta_k+1 (plaquette sums P_a, P_b)
leaves the product of their Boltzmann laws invariant iff the swap is
accepted with probability min(1, exp((beta_k+1 - beta_k) J (P_a - P_b))), the
weight ratio exp(beta_k P_b + beta_k+1 P_a) / exp(beta_k P_a + beta_k+1 P_b);
`check_exchange` compares a recorded chain with those laws.
"""
from __future__ import annotations

import math

import numpy as np


def plaquette_links_2d(L0: int, L1: int):
    """The four links of every plaquette of a 2D periodic plane, in the
    link numbering link = site * 2 + dir, site = x0 + L0 * x1
    (lattice.rs:259-263, 420-430)."""
    out = []
    for x1 in range(L1):
        for x0 in range(L0):
            s = x0 + L0 * x1
            s_right = (x0 + 1) % L0 + L0 * x1
            s_up = x0 + L0 * ((x1 + 1) % L1)
            out.append((2 * s, 2 * s_right + 1, 2 * s_up, 2 * s + 1))
    return out


def plaquette_dos(L0: int, L1: int):
    """(values, counts): how many of the 2^(2 L0 L1) link states have each
    signed plaquette sum."""
    nl = 2 * L0 * L1
    assert nl <= 20, "enumeration only for tiny planes"
    bits = (np.arange(1 << nl, dtype=np.int64)[:, None] >> np.arange(nl)[None, :]) & 1
    u = (1 - 2 * bits).astype(np.int8)
    ps = np.zeros(1 << nl, np.int64)
    for a, b, c, d in plaquette_links_2d(L0, L1):
        ps += (u[:, a] * u[:, b] * u[:, c] * u[:, d]).astype(np.int64)
    vals, counts = np.unique(ps, return_counts=True)
    return vals, counts.astype(np.float64)


def plaquette_law(dos, beta: float, j: float = 1.0) -> np.ndarray:
    """P(plaquette sum = vals[i]) at beta: weight exp(beta J P)."""
    vals, counts = dos
    logw = np.log(counts) + beta * j * vals
    w = np.exp(logw - logw.max())
    return w / w.sum()


def mean_plaquette_sum(dos, beta: float, j: float = 1.0) -> float:
    return float((plaquette_law(dos, beta, j) * dos[0]).sum())


def exact_swap_rate(dos, beta_lo: float, beta_hi: float, j: float = 1.0) -> float:
    """<min(1, exp((beta_hi - beta_lo) J (P_a - P_b)))> with P_a drawn at
    beta_lo and P_b at beta_hi, independently."""
    vals = dos[0]
    pa = plaquette_law(dos, beta_lo, j)
    pb = plaquette_law(dos, beta_hi, j)
    delta = (beta_hi - beta_lo) * j * (vals[:, None] - vals[None, :])
    acc = np.minimum(1.0, np.exp(np.minimum(delta, 0.0)))
    return float((pa[:, None] * pb[None, :] * acc).sum())


def check_exchange(sums, slots, betas, dos, reduce, burn=200, j=1.0, n_sigma=5.0):
    """Check a recorded exchange chain (one record per sweep, an exchange
    round after every record) of replicas of one shape.

    sums[t, r], slots[t, r]: replica r's plaquette sum and beta slot at record
    t, taken before round t.  betas must be ascending, so slot k is the k-th
    smallest beta.  `reduce` is the tau_int estimator (statistics.rs:83-137).
    Checks, each failing for a plausible mistake:
      1. every slot samples the exact law: <P> within n_sigma (a flipped
         delta sign or a swap of configurations instead of slots biases it);
      2. round t changes the assignment only by swapping the pairs of rank
         k = t mod 2, t mod 2 + 2, ... (a wrong pair parity fails);
      3. the swaps seen equal, in total, the acceptance probabilities
         min(1, e^delta) of the recorded sums within n_sigma binomial
         errors (a wrong threshold or comparison direction fails);
      4. the mean acceptance probability of each pair matches the exact
         <min(1, e^delta)> within n_sigma.
    Returns the per-pair (observed swap rate, exact rate)."""
    sums = np.asarray(sums, np.int64)
    slots = np.asarray(slots, np.int64)
    T, R = sums.shape
    assert list(betas) == sorted(betas)
    for k in range(R):
        series = np.array([sums[t][slots[t] == k][0] for t in range(burn, T)], np.float64)
        est = reduce(series)
        want = mean_plaquette_sum(dos, betas[k], j)
        assert abs(est.mean - want) < n_sigma * est.stderr + 1e-12, (k, est.mean, want, est.stderr)
    p_all = [[] for _ in range(R - 1)]
    s_all = [[] for _ in range(R - 1)]
    for t in range(T - 1):
        rep_at = np.empty(R, np.int64)
        rep_at[slots[t]] = np.arange(R)
        nxt = rep_at.copy()
        for k in range(t % 2, R - 1, 2):
            a, b = rep_at[k], rep_at[k + 1]
            delta = (betas[k + 1] - betas[k]) * j * float(sums[t, a] - sums[t, b])
            p = 1.0 if delta >= 0 else math.exp(delta)
            swapped = slots[t + 1, a] == k + 1
            if swapped:
                nxt[k], nxt[k + 1] = b, a
            if t >= burn:
                p_all[k].append(p)
                s_all[k].append(1.0 if swapped else 0.0)
        want_rep = np.empty(R, np.int64)
        want_rep[slots[t + 1]] = np.arange(R)
        assert np.array_equal(want_rep, nxt), f"round {t} changed pairs of the other parity"
    out = []
    for k in range(R - 1):
        p = np.array(p_all[k])
        s = np.array(s_all[k])
        assert p.size > 100
        var = float((p * (1 - p)).sum())
        assert abs(s.sum() - p.sum()) < n_sigma * math.sqrt(var) + 1e-9, (k, s.sum(), p.sum(), var)
        est = reduce(p)
        want = exact_swap_rate(dos, betas[k], betas[k + 1], j)
        assert abs(est.mean - want) < n_sigma * est.stderr + 1e-12, (k, est.mean, want, est.stderr)
        out.append((float(s.mean()), want))
    return out
