"""Pins of the oracle's per-configuration measurements (CPU only).

The correlator pair sums, Polyakov loops and Wilson rectangles of
oracle/plq_oracle_measure.c are checked against the known answers of the
reference's own unit tests (ising/mod.rs:211-246, potts/mod.rs:501-540,
gauge/mod.rs:645-792), against closed forms, and against independent numpy
computations: a plane-wave configuration's correlator, the enclosed-plaquette
identity of every Wilson rectangle, gauge invariance and the centre flip.
"""
import numpy as np
import pytest

import oracle


def _grid(shape, state):
    """Lattice-order states as a numpy grid indexed [x_{D-1}, ..., x_0]."""
    return np.asarray(state).reshape(tuple(reversed(shape)))


def _axis(shape, mu):
    return len(shape) - 1 - mu


# ----------------------------------------------------------- correlator ---

def test_aligned_ising_correlator_is_all_ones():
    c = oracle.ising_correlator([4, 4], np.zeros(16, np.uint8))
    assert len(c) == 2
    for row in c:
        assert row.tolist() == [1.0, 1.0, 1.0]


def test_checkerboard_ising_correlator_alternates():
    st = np.array([(x + y) % 2 for y in range(4) for x in range(4)], np.uint8)
    c = oracle.ising_correlator([4, 4], st)
    assert c[0].tolist() == [1.0, -1.0, 1.0]
    assert c[1].tolist() == [1.0, -1.0, 1.0]


def test_uniform_potts_correlator_is_one_minus_one_over_q():
    q = 3
    c = oracle.potts_correlator([4, 6], q, np.zeros(24, np.uint8))
    for mu, row in enumerate(c):
        assert len(row) == [4, 6][mu] // 2 + 1
        assert np.all(row == 1.0 - 1.0 / q)


def test_independent_potts_labels_correlate_near_zero():
    q = 3
    rng = np.random.default_rng(20260809)
    c = oracle.potts_correlator([32, 32], q, rng.integers(0, q, 1024, dtype=np.uint8))
    for row in c:
        assert abs(row[0] - (1.0 - 1.0 / q)) < 1e-12
        assert np.all(np.abs(row[1:]) < 0.05)


@pytest.mark.parametrize("shape", [[6, 4, 5], [7, 3], [9], [4, 3, 2, 5]])
def test_plane_wave_correlator_matches_its_one_dimensional_closed_form(shape):
    # States varying along one axis only: every other axis correlates fully,
    # and along that axis C_r is the 1D autocorrelation of the profile.
    rng = np.random.default_rng(len(shape) * 17 + shape[0])
    for mu in range(len(shape)):
        prof = rng.integers(0, 2, shape[mu], dtype=np.uint8)
        coords = np.indices(tuple(reversed(shape)))[_axis(shape, mu)]
        st = prof[coords].reshape(-1)
        sums = oracle.axis_pair_sums(shape, st, 0)
        n = int(np.prod(shape))
        sign = 1 - 2 * prof.astype(np.int64)
        for nu, row in enumerate(sums):
            if nu != mu:
                assert row.tolist() == [n] * (shape[nu] // 2 + 1)
        want = [int((sign * np.roll(sign, -r)).sum()) * (n // shape[mu]) for r in range(shape[mu] // 2 + 1)]
        assert sums[mu].tolist() == want


@pytest.mark.parametrize("shape,q", [([5, 4, 3], 2), ([6, 7], 4), ([3, 3, 3, 4], 3), ([11], 5)])
def test_pair_sums_match_shifted_grids(shape, q):
    rng = np.random.default_rng(sum(shape) * q)
    st = rng.integers(0, q, int(np.prod(shape)), dtype=np.uint8)
    g = _grid(shape, st)
    kind = 0 if q == 2 else 1
    sums = oracle.axis_pair_sums(shape, st, kind)
    for mu in range(len(shape)):
        for r in range(shape[mu] // 2 + 1):
            other = np.roll(g, -r, axis=_axis(shape, mu))
            want = ((1 - 2 * g.astype(np.int64)) * (1 - 2 * other.astype(np.int64))).sum() if kind == 0 \
                else (g == other).sum()
            assert sums[mu][r] == want, (mu, r)


# ------------------------------------------------------------- gauge ---

def _hot_links(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2, int(np.prod(shape)) * len(shape), dtype=np.uint8)


def _gauge_transform(shape, links, sites):
    """Flip every link touching each site (lattice.rs:373-385)."""
    D = len(shape)
    out = links.copy()
    stride = np.cumprod([1] + list(shape[:-1]))
    for s in sites:
        for mu in range(D):
            out[s * D + mu] ^= 1
            x = (s // stride[mu]) % shape[mu]
            back = s - x * stride[mu] + ((x - 1) % shape[mu]) * stride[mu]
            out[back * D + mu] ^= 1
    return out


def _plane_plaquettes(shape, links, mu, nu):
    """U_mu(x) U_nu(x + e_mu) U_mu(x + e_nu) U_nu(x) as a +-1 grid."""
    D = len(shape)
    u = [(1 - 2 * _grid(shape, links[a::D]).astype(np.int64)) for a in range(D)]
    am, an = _axis(shape, mu), _axis(shape, nu)
    return u[mu] * np.roll(u[nu], -1, axis=am) * np.roll(u[mu], -1, axis=an) * u[nu]


def _enclosed_sum(shape, links, r, t):
    """Sum over ordered pairs and sites of the product of the enclosed plaquettes."""
    D = len(shape)
    total = 0
    for mu in range(D):
        for nu in range(D):
            if mu == nu:
                continue
            p = _plane_plaquettes(shape, links, mu, nu)
            acc = np.ones_like(p)
            for a in range(r):
                for b in range(t):
                    acc = acc * np.roll(np.roll(p, -a, axis=_axis(shape, mu)), -b, axis=_axis(shape, nu))
            total += int(acc.sum())
    return total


def test_cold_polyakov_loop_is_one_in_every_direction():
    shape = [4, 5, 6]
    links = np.zeros(120 * 3, np.uint8)
    for d in range(3):
        assert oracle.polyakov_loop(shape, links, d) == 1.0


def test_centre_transformation_flips_one_polyakov_direction():
    shape = [4, 4, 4]
    links = _hot_links(shape, 13)
    before = [oracle.polyakov_loop(shape, links, d) for d in range(3)]
    flipped = links.copy()
    for s in range(64):
        if s % 4 == 2:
            flipped[s * 3 + 0] ^= 1
    assert oracle.plaquette_sum(shape, flipped) == oracle.plaquette_sum(shape, links)
    assert oracle.polyakov_loop(shape, flipped, 0) == -before[0]
    assert oracle.polyakov_loop(shape, flipped, 1) == before[1]
    assert oracle.polyakov_loop(shape, flipped, 2) == before[2]


def test_one_flipped_link_flips_exactly_its_line():
    shape = [3, 4, 5]
    links = _hot_links(shape, 29)
    D, n = 3, 60
    stride = [1, 3, 12]
    for site, d in [(0, 0), (17, 1), (59, 2), (31, 1)]:
        other = links.copy()
        other[site * D + d] ^= 1
        # The line through `site` along d: its product is that of the original.
        base = site - ((site // stride[d]) % shape[d]) * stride[d]
        prod = 1
        for k in range(shape[d]):
            prod *= 1 - 2 * int(links[(base + k * stride[d]) * D + d])
        lines = n // shape[d]
        assert oracle.polyakov_loop(shape, other, d) == (oracle.polyakov_loop(shape, links, d) * lines - 2 * prod) / lines
        for e in range(D):
            if e != d:
                assert oracle.polyakov_loop(shape, other, e) == oracle.polyakov_loop(shape, links, e)


def test_gauge_transformation_leaves_polyakov_and_wilson_unchanged():
    shape = [4, 4, 4]
    links = _hot_links(shape, 17)
    moved = _gauge_transform(shape, links, range(0, 64, 5))
    assert not np.array_equal(moved, links)
    for d in range(3):
        assert oracle.polyakov_loop(shape, moved, d) == oracle.polyakov_loop(shape, links, d)
    assert np.array_equal(oracle.wilson_rectangles(shape, moved, 2), oracle.wilson_rectangles(shape, links, 2))


def test_cold_wilson_table_is_all_ones():
    table = oracle.wilson_rectangles([6, 6, 6], np.zeros(216 * 3, np.uint8), 3)
    assert table.shape == (4, 4)
    assert np.all(table == 1.0)


@pytest.mark.parametrize("shape", [[6, 6], [4, 4, 4]])
def test_unit_wilson_loop_is_the_mean_plaquette(shape):
    links = _hot_links(shape, 11)
    D, n = len(shape), int(np.prod(shape))
    mean = oracle.plaquette_sum(shape, links) / (n * D * (D - 1) // 2)
    assert oracle.wilson_rectangles(shape, links, 2)[1][1] == mean


def test_wilson_table_is_symmetric_and_capped():
    table = oracle.wilson_rectangles([4, 8], _hot_links([4, 8], 3), 10)
    assert table.shape == (3, 3)
    assert np.array_equal(table, table.T)


@pytest.mark.parametrize("shape,side", [([8, 6], 3), ([5, 4, 6], 2), ([4, 4, 2, 4], 1), ([6, 4, 4, 4], 2)])
def test_every_rectangle_is_the_product_of_its_enclosed_plaquettes(shape, side):
    links = _hot_links(shape, sum(shape) + side)
    totals, placements = oracle.wilson_sums(shape, links, side)
    D = len(shape)
    assert placements == D * (D - 1) * int(np.prod(shape))
    for r in range(1, side + 1):
        for t in range(1, side + 1):
            assert totals[r][t] == _enclosed_sum(shape, links, r, t), (r, t)


# ------------------------------------------- 1D checkerboard Metropolis ---

@pytest.mark.parametrize("seed", [1, 2, 3])
def test_one_dimensional_checkerboard_metropolis_conserves_the_wall_imbalance(seed):
    # At h = 0 a site with one domain wall beside it has dE = 0 and always
    # flips, so under the two-colour order every isolated wall moves two
    # bonds per sweep: walls on even bonds (i, i+1), i even, one way, walls on
    # odd bonds the other, and walls are made and destroyed one of each kind.
    # The difference of the two counts is conserved, the chain is not ergodic,
    # and its correlator misses the exact 1D result; the heat bath has no such
    # invariant.
    N = 16
    st = np.random.default_rng(seed).integers(0, 2, N).astype(np.uint8)

    def imbalance(s):
        w = s != np.roll(s, -1)
        return int(w[0::2].sum()) - int(w[1::2].sum())

    for kernel, conserved in [(oracle.KERNEL_METROPOLIS, True), (oracle.KERNEL_HEAT_BATH, False)]:
        d = oracle.local_desc([N], model=oracle.MODEL_ISING, q=2, kernel=kernel, beta=0.4, j=1.0, h=0.0, seed=seed)
        s, seen = st.copy(), set()
        for sweep in range(400):
            s = oracle.local_sweeps(d, s, sweep, 1)
            seen.add(imbalance(s))
        assert (seen == {imbalance(st)}) == conserved, (kernel, seen)


@pytest.mark.parametrize("seed", [3, 4])
def test_two_dimensional_link_checkerboard_metropolis_conserves_the_vortex_imbalance(seed):
    # 2D Z2: a link between a negative and a positive plaquette has dE = 0 and
    # always flips, moving the negative plaquette (a vortex) across it.  In
    # each of the four passes of a sweep (direction-major, colour 0 then 1)
    # every isolated vortex crosses the one link of that pass's colour on its
    # boundary and so changes plaquette parity, and vortices are made and
    # destroyed in adjacent pairs of opposite parity.  The difference of the
    # vortex counts on even and odd plaquettes is therefore the same after
    # every sweep under Metropolis, and free under the heat bath.
    shape = [8, 8]
    st = np.random.default_rng(seed).integers(0, 2, 128).astype(np.uint8)
    parity = np.indices((8, 8)).sum(axis=0) & 1

    def imbalance(s):
        neg = _plane_plaquettes(shape, s, 0, 1) < 0
        return int(neg[parity == 0].sum()) - int(neg[parity == 1].sum())

    for kernel, conserved in [(oracle.KERNEL_METROPOLIS, True), (oracle.KERNEL_HEAT_BATH, False)]:
        d = oracle.local_desc(shape, model=oracle.MODEL_Z2, q=2, kernel=kernel, beta=0.5, j=1.0, h=0.0, seed=seed)
        s, seen = st.copy(), set()
        for sweep in range(200):
            s = oracle.local_sweeps(d, s, sweep, 1)
            seen.add(imbalance(s))
        assert (seen == {imbalance(st)}) == conserved, (kernel, seen)
