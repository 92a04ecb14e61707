"""The product's FP64 statistics (plaquette_amd.statistics) against the
reference's golden numbers (statistics.rs:613-641) and the oracle's
restatement of statistics.rs:83-317, on the direct-sum and FFT paths."""
import math

import numpy as np
import pytest

import oracle as O
from plaquette_amd import statistics as S


M64 = (1 << 64) - 1


def base_value(i):
    # statistics.rs:343-349 (splitmix64 finaliser on i)
    z = (i + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    z ^= z >> 31
    return (z >> 11) / float(1 << 53) * 2.0 - 1.0


def test_golden_numbers():
    # statistics.rs:613-641, relative tolerance 1e-9
    beta, n = 0.4407, 64.0
    signed = np.array([(1.0 if (i // 500) % 2 == 0 else -1.0) * n * (0.6 + 0.15 * base_value(i // 3))
                       for i in range(3000)])
    e = -1.9 * n + 0.3 * np.abs(signed)

    def close(got, want):
        assert abs(got - want) <= 1e-9 * abs(want), (got, want)

    c = S.specific_heat(e, beta, n)
    close(c.value, 8.24400144960567e-3)
    close(c.stderr, 2.377913272032474e-4)
    assert c.n_blocks == 1000
    chi = S.susceptibility(signed, beta, n)
    close(chi.value, 2.078511824523012e-1)
    close(chi.stderr, 5.437844369399705e-3)
    assert chi.n_blocks == 750
    u4 = S.binder(signed)
    close(u4.value, 6.403580348646347e-1)
    close(u4.stderr, 7.451725532088574e-4)
    assert u4.n_blocks == 750


def _ar1(n, phi, seed):
    rng = np.random.default_rng(seed)
    x = np.empty(n)
    x[0] = rng.standard_normal()
    eps = rng.standard_normal(n)
    for i in range(1, n):
        x[i] = phi * x[i - 1] + eps[i]
    return x


@pytest.mark.parametrize("n,phi", [(500, 0.5), (3000, 0.9), (5000, 0.95), (20000, 0.99)])
def test_reduce_matches_the_oracle(n, phi):
    # n >= 4096 takes the FFT path
    x = _ar1(n, phi, n)
    got, want = S.reduce(x), O.reduce(x)
    assert abs(got.mean - want.mean) <= 1e-12 * max(1.0, abs(want.mean))
    for a, b in ((got.tau_int, want.tau_int), (got.stderr, want.stderr), (got.n_eff, want.n_eff)):
        assert abs(a - b) <= 1e-9 * abs(b), (a, b)


def test_ar1_tau_int_closed_form():
    # AR(1): tau_int = (1 + phi) / (2 (1 - phi)) = 9.5 at phi = 0.9
    est = S.reduce(_ar1(400000, 0.9, 7))
    assert abs(est.tau_int - 9.5) < 0.5


def test_reduce_limits():
    # statistics.rs:368-394
    alt = np.array([1.0 if i % 2 == 0 else -1.0 for i in range(512)])
    est = S.reduce(alt)
    assert abs(est.tau_int - 0.5) < 1e-9 and abs(est.n_eff - 512) < 1e-9
    est = S.reduce(np.full(128, 3.0))
    assert est.stderr == 0.0 and est.tau_int == 0.5 and est.n_eff == 128.0 and est.mean == 3.0


def test_binder_limits_and_short_series():
    # statistics.rs:674-680, 771-781, 444-451
    assert abs(S.binder(np.full(100, 2.0)).value - 2.0 / 3.0) < 1e-12
    m = np.zeros(300)
    m[:100] = 1.0
    assert abs(S.binder(m).value) < 1e-12
    d = S.susceptibility(np.array([48.0]), 0.4407, 64.0)
    assert d.n_blocks < 2 and math.isnan(d.stderr) and math.isfinite(d.value)
