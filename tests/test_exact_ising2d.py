"""Kaufman's finite-torus partition function (tests/exact_ising2d.py) against
brute-force enumeration, so it can pin the samplers at sizes no enumeration
reaches (2D Ising 64^2 at beta_c)."""
import math

import numpy as np
import pytest

import exact_ising2d as X


@pytest.mark.parametrize("m,n", [(4, 4), (2, 4), (4, 2), (2, 6), (6, 2)])
@pytest.mark.parametrize("beta", [0.1, 0.3, 0.4406868, 0.7, 1.2])
def test_kaufman_matches_enumeration(m, n, beta):
    lz, e = X.enumerate_log_z(m, n, beta)
    assert abs(X.log_z(m, n, beta) - lz) < 1e-12 * abs(lz)
    assert abs(X.energy_per_site(m, n, beta) - e) < 1e-7


def test_specific_heat_matches_enumeration():
    m, n, beta = 4, 4, 0.4406868
    h = 1e-4
    lz = [X.enumerate_log_z(m, n, beta + d)[0] for d in (-h, 0.0, h)]
    c = beta * beta * (lz[0] - 2 * lz[1] + lz[2]) / (h * h) / (m * n)
    assert abs(X.specific_heat(m, n, beta) - c) < 1e-5


def test_large_torus_approaches_onsager():
    # E/N -> -coth(2 beta_c) (1 + ...) = -sqrt(2) at beta_c; 512^2 is within 1e-3
    assert abs(X.energy_per_site(512, 512, 0.4406868) + math.sqrt(2)) < 2e-3
    # deep in either phase the finite torus matches the infinite lattice
    assert abs(X.energy_per_site(64, 64, 0.1) - X.energy_per_site(128, 128, 0.1)) < 1e-9
