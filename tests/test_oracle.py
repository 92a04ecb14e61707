"""Pin the CPU oracle before trusting it (CPU only).

Each check comes from something the oracle's formula does not restate: the
reference's own known-answer tests and golden numbers (cited file:line), exact
enumeration, closed forms, invariants, or an independently derived inverse.
"""
import ctypes as C
import itertools
import math

import numpy as np
import pytest

import oracle as O

# ----------------------------------------------------------------- RNG ------

M32 = 0xFFFFFFFF


def _inv_xorshift(y, s):
    # inverse of x ^= x >> s on 32 bits
    x = y
    for _ in range(32 // s + 1):
        x = y ^ (x >> s)
    return x & M32


def _inv_lowbias32(y):
    # The published lowbias32 (rng.wgsl:34-42) is five invertible steps; undo
    # them with the multiplicative inverses of its two constants.
    inv1 = pow(0x7FEB352D, -1, 1 << 32)
    inv2 = pow(0x846CA68B, -1, 1 << 32)
    x = _inv_xorshift(y, 16)
    x = (x * inv2) & M32
    x = _inv_xorshift(x, 15)
    x = (x * inv1) & M32
    x = _inv_xorshift(x, 16)
    return x


def test_lowbias32_is_the_published_bijection():
    rng = np.random.default_rng(1)
    for x in [0, 1, 2, 0xFFFFFFFF, 0x80000000] + [int(v) for v in rng.integers(0, 1 << 32, 2000)]:
        assert _inv_lowbias32(O.lowbias32(x)) == x
    assert O.lowbias32(0) == 0  # every step fixes zero


def test_keyed_draw_composes_two_hashes_and_is_uniform():
    # keyed(seed, var, sweep) = lowbias32(lowbias32(seed ^ var) ^ sweep*0x9e3779b9) >> 8
    # (rng.wgsl:47-52).  The draw keeps the top 24 bits of the outer hash, so
    # try all 256 values of the dropped low byte: undo the outer hash with the
    # independent inverse, remove the sweep key, undo the inner hash, and
    # exactly the true seed ^ var must come back for some completion.  A
    # missing or swapped round, a wrong sweep constant or a wrong shift fails.
    seed = 0xDEADBEEF
    cases = [(0, 0), (5, 0), (12345, 7), (1 << 20, 999), (0xFFFFFFFF, 1), (77, 0xFFFFFFFF)]
    rng = np.random.default_rng(3)
    cases += [(int(v), int(s)) for v, s in zip(rng.integers(0, 1 << 32, 40), rng.integers(0, 1 << 32, 40))]
    for var, sweep in cases:
        m = O.keyed_bits(seed, var, sweep)
        assert 0 <= m < (1 << 24)
        assert O.keyed_uniform(seed, var, sweep) == m / 16777216.0
        key = (sweep * 0x9E3779B9) & M32
        inner = [_inv_lowbias32(_inv_lowbias32((m << 8) | low) ^ key) for low in range(256)]
        assert (seed ^ var) & M32 in inner, (var, sweep)
        # every completion is an input that hashes forward to the same draw
        for x in inner:
            assert O.keyed_bits(0, x, sweep) == m
    assert O.keyed_bits(seed, 9, 3) != O.keyed_bits(seed, 9, 4)
    # low-order chi-square over 64 bins of many draws
    ms = np.array([O.keyed_bits(7, v, 3) for v in range(20000)])
    counts = np.bincount(ms >> 18, minlength=64)
    chi2 = ((counts - 20000 / 64) ** 2 / (20000 / 64)).sum()
    assert chi2 < 120  # 63 dof; p ~ 1e-5 bound


def test_fold_seed_mixes_the_high_word():
    # device.rs:761-763
    assert O.fold_seed(0x0000000100000000) == 1
    assert O.fold_seed(0xFFFFFFFF00000000) == 0xFFFFFFFF
    assert O.fold_seed(12345) == 12345


# ------------------------------------------------------------- geometry -----

def test_known_neighbors_on_3x3():
    # lattice.rs:595-603
    nb = O.neighbor_table([3, 3])
    assert list(nb[4]) == [5, 3, 7, 1]
    assert list(nb[0]) == [1, 2, 3, 6]


def test_generalizes_to_three_dimensions():
    # lattice.rs:605-613
    nb = O.neighbor_table([2, 3, 4]).ravel()
    assert nb.size == 24 * 6
    assert nb[4] == 6 and nb[5] == 18


def test_parity_is_the_coordinate_sum():
    # lattice.rs:585-593
    for shape in ([4, 6, 2], [3, 5, 3]):
        par = O.parity_table(shape)
        for site in range(int(np.prod(shape))):
            rest, total = site, 0
            for l in shape:
                total += rest % l
                rest //= l
            assert par[site] == total % 2


def test_hand_computed_plaquettes_on_3x3x3():
    # lattice.rs:674-684
    assert O.plaquette_links([3, 3, 3], 0) == [0, 4, 9, 1]
    assert O.plaquette_links([3, 3, 3], 2) == [1, 11, 28, 2]
    assert O.plaquette_links([3, 3, 3], 4 * 3 + 1) == [12, 17, 39, 14]


@pytest.mark.parametrize("shape", [[4, 4], [2, 3, 4], [3, 3, 3], [2, 2, 2, 2]])
def test_staple_groups_are_the_plaquettes_containing_the_link(shape):
    # lattice.rs:717-790: groups in increasing plaquette order, each the
    # plaquette's other three links
    d = len(shape)
    n = int(np.prod(shape))
    n_plaq = n * d * (d - 1) // 2
    containing = {}
    for p in range(n_plaq):
        for link in O.plaquette_links(shape, p):
            containing.setdefault(link, []).append(p)
    st = O.staple_table(shape)
    for link in range(n * d):
        ps = containing[link]
        assert len(ps) == 2 * (d - 1)
        for g, p in enumerate(ps):
            want = [l for l in O.plaquette_links(shape, p) if l != link]
            assert list(st[link, 3 * g:3 * g + 3]) == want


# ------------------------------------------------------------ statistics -----

M64 = (1 << 64) - 1


def base_value(i):
    # statistics.rs:343-349 (splitmix64 finaliser on i)
    z = (i + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    z ^= z >> 31
    return (z >> 11) / float(1 << 53) * 2.0 - 1.0


def reference_series():
    beta, n = 0.4407, 64.0
    signed = np.array([(1.0 if (i // 500) % 2 == 0 else -1.0) * n * (0.6 + 0.15 * base_value(i // 3))
                       for i in range(3000)])
    e = -1.9 * n + 0.3 * np.abs(signed)
    return e, signed, beta, n


def test_statistics_golden_numbers():
    # statistics.rs:613-641, relative tolerance 1e-9
    e, signed, beta, n = reference_series()

    def close(got, want):
        assert abs(got - want) <= 1e-9 * abs(want)

    c = O.specific_heat(e, beta, n)
    close(c.value, 8.24400144960567e-3)
    close(c.stderr, 2.377913272032474e-4)
    assert c.n_blocks == 1000
    chi = O.susceptibility(signed, beta, n)
    close(chi.value, 2.078511824523012e-1)
    close(chi.stderr, 5.437844369399705e-3)
    assert chi.n_blocks == 750
    u4 = O.binder(signed)
    close(u4.value, 6.403580348646347e-1)
    close(u4.stderr, 7.451725532088574e-4)
    assert u4.n_blocks == 750


def test_reduce_limits():
    # statistics.rs:368-394
    alt = np.array([1.0 if i % 2 == 0 else -1.0 for i in range(512)])
    est = O.reduce(alt)
    assert abs(est.tau_int - 0.5) < 1e-9 and abs(est.n_eff - 512) < 1e-9
    est = O.reduce(np.full(128, 3.0))
    assert est.stderr == 0.0 and est.tau_int == 0.5 and est.n_eff == 128.0 and est.mean == 3.0
    # tau_int rises with run length (statistics.rs:397-418)
    taus = [O.reduce(np.array([base_value(i // k) for i in range(1024)])).tau_int for k in (1, 4, 16)]
    assert taus[0] < taus[1] < taus[2]


def test_binder_limits():
    # statistics.rs:674-680 and 771-781
    assert abs(O.binder(np.full(100, 2.0)).value - 2.0 / 3.0) < 1e-9
    m = np.zeros(300)
    m[:100] = 1.0
    assert abs(O.binder(m).value) < 1e-9


def test_jackknife_refuses_short_series():
    # statistics.rs:444-451, 645-653
    d = O.susceptibility(np.array([48.0]), 0.4407, 64.0)
    assert d.n_blocks < 2 and math.isnan(d.stderr) and math.isfinite(d.value)


# ------------------------------------------------------ local updates -------

def test_zero_coupling_negative_field_flips_every_site_once():
    # ising/gpu.rs:159-192: J = 0, h = -1 from a cold start flips every site
    # on its first visit whatever the neighbours do, so M = -N after a sweep
    for shape in ([8, 8], [4, 6, 8], [4, 4, 4, 2]):
        n = int(np.prod(shape))
        d = O.local_desc(shape, beta=1.0, j=0.0, h=-1.0, seed=7)
        st = O.local_sweeps(d, np.zeros(n, np.uint8), 0, 1)
        b, s = O.ising_sums(shape, st)
        assert s == -n


def _exact_ising(shape, beta, j=1.0, h=0.0):
    n = int(np.prod(shape))
    nb = O.neighbor_table(shape)
    fwd = nb[:, 0::2]
    Z = 0.0
    E1 = 0.0
    M2 = 0.0
    for bits in range(1 << n):
        s = np.array([1 - 2 * ((bits >> k) & 1) for k in range(n)])
        e = -j * float((s[:, None] * s[fwd]).sum()) - h * float(s.sum())
        w = math.exp(-beta * e)
        Z += w
        E1 += w * e
        M2 += w * float(s.sum()) ** 2
    return E1 / Z, M2 / Z


@pytest.mark.parametrize("kernel", [O.KERNEL_METROPOLIS, O.KERNEL_HEAT_BATH])
@pytest.mark.parametrize("profile", [O.PROFILE_F64, O.PROFILE_F32])
def test_checkerboard_samples_the_exact_4x4_distribution(kernel, profile):
    # Exact enumeration of the 2^16 states of a 4x4 torus; a sign slip in the
    # energy delta, a wrong neighbour or a wrong acceptance rule moves <E>
    # by many standard errors.
    shape, beta = [4, 4], 0.35
    e_exact, m2_exact = _exact_ising(shape, beta)
    d = O.local_desc(shape, kernel=kernel, profile=profile, beta=beta, seed=2024)
    st = np.zeros(16, np.uint8)
    es, m2s = [], []
    for t in range(20000):
        st = O.local_sweeps(d, st, t, 1)
        b, s = O.ising_sums(shape, st)
        if t >= 200:
            es.append(-float(b))
            m2s.append(float(s) ** 2)
    for series, exact in ((es, e_exact), (m2s, m2_exact)):
        est = O.reduce(np.array(series))
        assert abs(est.mean - exact) < 5 * est.stderr + 1e-12, (est.mean, exact, est.stderr)


def test_potts_q2_at_twice_the_coupling_is_ising_in_lockstep():
    # potts_e2e.rs:758-812: Potts(q=2, 2J) and Ising(J) are one chain; with
    # the same keyed draws the two model paths must agree state for state
    shape = [6, 4]
    rng = np.random.default_rng(3)
    start = rng.integers(0, 2, 24).astype(np.uint8)
    for kernel in (O.KERNEL_METROPOLIS, O.KERNEL_HEAT_BATH):
        di = O.local_desc(shape, model=O.MODEL_ISING, kernel=kernel, beta=0.44, j=1.0, seed=9)
        dp = O.local_desc(shape, model=O.MODEL_POTTS, q=2, kernel=kernel, beta=0.44, j=2.0, seed=9)
        a = O.local_sweeps(di, start, 0, 50)
        b = O.local_sweeps(dp, start, 0, 50)
        assert np.array_equal(a, b)


def test_metropolis_limits():
    # beta = 0 accepts every flip (a checkerboard sweep then flips all sites);
    # a huge beta never accepts an uphill flip, so a cold start stays put
    shape = [6, 6]
    d0 = O.local_desc(shape, beta=0.0, seed=1)
    st = O.local_sweeps(d0, np.zeros(36, np.uint8), 0, 1)
    assert st.sum() == 36
    dinf = O.local_desc(shape, beta=50.0, seed=1)
    st = O.local_sweeps(dinf, np.zeros(36, np.uint8), 0, 20)
    assert st.sum() == 0


def test_heat_bath_at_zero_coupling_exposes_the_draw():
    # J = h = 0: both weights are 1, so the new state is 0 iff u * 2 < 1,
    # i.e. iff bit 23 of the keyed draw is clear
    shape = [4, 6]
    d = O.local_desc(shape, kernel=O.KERNEL_HEAT_BATH, beta=0.7, j=0.0, h=0.0, seed=77)
    st = O.local_sweeps(d, np.zeros(24, np.uint8), 5, 1)
    seed32 = O.fold_seed(77)
    want = np.array([(O.keyed_bits(seed32, s, 5) >> 23) & 1 for s in range(24)], np.uint8)
    assert np.array_equal(st, want)


def _exact_z2_plaquette(shape, beta):
    n = int(np.prod(shape))
    d = len(shape)
    nl = n * d
    plaqs = [O.plaquette_links(shape, p) for p in range(n * d * (d - 1) // 2)]
    Z = P = 0.0
    for bits in range(1 << nl):
        u = [1 - 2 * ((bits >> k) & 1) for k in range(nl)]
        ps = sum(u[a] * u[b] * u[c] * u[e] for a, b, c, e in plaqs)
        w = math.exp(beta * ps)
        Z += w
        P += w * ps
    return P / Z


def test_z2_link_checkerboard_heat_bath_samples_the_exact_distribution():
    # A 2x4 plane has 16 links: enumerate all 2^16 states.  (Metropolis is
    # not used here: with its deterministic flip, the two-dimensional link
    # checkerboard order of updater.rs:325-343 is not ergodic on small planes
    # -- the same bias appears with independent uniforms in place of the keyed
    # draws -- so it cannot be checked against the Boltzmann weights there.)
    shape, beta = [2, 4], 0.6
    exact = _exact_z2_plaquette(shape, beta)
    d = O.local_desc(shape, model=O.MODEL_Z2, kernel=O.KERNEL_HEAT_BATH, beta=beta, seed=5)
    st = np.zeros(16, np.uint8)
    ps = []
    for t in range(20000):
        st = O.local_sweeps(d, st, t, 1)
        if t >= 200:
            ps.append(float(O.plaquette_sum(shape, st)))
    est = O.reduce(np.array(ps))
    assert abs(est.mean - exact) < 5 * est.stderr, (est.mean, exact, est.stderr)


def test_z2_three_dimensional_metropolis_agrees_with_heat_bath():
    # gauge_e2e.rs:355-404: the link checkerboard measures the same physics
    # under either kernel; 4 sigma on the tau_int-inflated errors
    shape, beta = [4, 4, 4], 0.6
    means = []
    for kernel in (O.KERNEL_METROPOLIS, O.KERNEL_HEAT_BATH):
        d = O.local_desc(shape, model=O.MODEL_Z2, kernel=kernel, beta=beta, seed=5 + kernel)
        st = np.zeros(192, np.uint8)
        ps = []
        for t in range(1500):
            st = O.local_sweeps(d, st, t, 1)
            if t >= 100:
                ps.append(O.plaquette_sum(shape, st) / 192.0)
        means.append(O.reduce(np.array(ps)))
    gap = abs(means[0].mean - means[1].mean)
    assert gap < 4 * math.hypot(means[0].stderr, means[1].stderr)


def test_z2_two_dimensional_strong_coupling_is_tanh_beta():
    # gauge_e2e.rs:243-263 / docs: in 2D the mean plaquette is tanh(beta)
    # up to finite-volume corrections; at 8x8 they are tiny
    shape, beta = [8, 8], 0.5
    d = O.local_desc(shape, model=O.MODEL_Z2, kernel=O.KERNEL_HEAT_BATH, beta=beta, seed=11)
    st = np.zeros(128, np.uint8)
    ps = []
    for t in range(6000):
        st = O.local_sweeps(d, st, t, 1)
        if t >= 200:
            ps.append(O.plaquette_sum(shape, st) / 64.0)
    est = O.reduce(np.array(ps))
    assert abs(est.mean - math.tanh(beta)) < 5 * est.stderr + 2e-3


# ------------------------------------------------------- replica exchange ---

def test_exact_z2_helper_matches_the_oracle_geometry():
    # the enumerated density of states uses its own plaquette list; pin it
    # against the oracle's plaquette_links (lattice.rs:420-430) and sums
    import exact_z2 as X

    dos = X.plaquette_dos(2, 4)
    assert dos[1].sum() == 1 << 16
    assert abs(X.mean_plaquette_sum(dos, 0.6) - _exact_z2_plaquette([2, 4], 0.6)) < 1e-9
    rng = np.random.default_rng(9)
    plaqs = X.plaquette_links_2d(2, 4)
    for _ in range(50):
        st = rng.integers(0, 2, 16).astype(np.uint8)
        u = 1 - 2 * st.astype(np.int64)
        assert sum(u[a] * u[b] * u[c] * u[d] for a, b, c, d in plaqs) == O.plaquette_sum([2, 4], st)


@pytest.mark.parametrize("betas", [[0.25, 0.55, 0.85], [0.3, 0.9]])
def test_replica_exchange_samples_the_exact_laws(betas):
    # Heat-bath replicas of a 2x4 Z2 plane (2^16 states, enumerated) with an
    # exchange round after every sweep: every beta slot must sample its exact
    # plaquette law, the swaps must follow min(1, e^delta) of the recorded
    # sums on the pairs of the round's parity, and each pair's mean
    # acceptance must equal the exact <min(1, e^delta)>.
    import exact_z2 as X

    R = len(betas)
    ens = O.Ensemble(2, [[2, 4]] * R, betas, [31 + r for r in range(R)], kernel=O.KERNEL_HEAT_BATH,
                     measure_every=1, exchange=True, exchange_seed=17)
    ens.advance(40000)
    sums, slots = np.array(ens.records), np.array(ens.slots)
    rates = X.check_exchange(sums, slots, betas, X.plaquette_dos(2, 4), O.reduce)
    # the oracle's own counters agree with the swaps seen in the slots
    seen = np.zeros(R - 1)
    for t in range(len(slots) - 1):
        rep_at = np.empty(R, np.int64)
        rep_at[slots[t]] = np.arange(R)
        for k in range(t % 2, R - 1, 2):
            seen[k] += slots[t + 1, rep_at[k]] == k + 1
    assert np.array_equal(seen, ens.accepted[:R - 1].astype(np.float64) - _last_round_swaps(ens, slots, R))
    assert all(0.05 < r[1] < 0.95 for r in rates)


def _last_round_swaps(ens, slots, R):
    # the final round's swaps happened after the last record was taken
    t = len(slots) - 1
    cur = ens.slot_of
    out = np.zeros(R - 1)
    rep_at = np.empty(R, np.int64)
    rep_at[slots[t]] = np.arange(R)
    for k in range(t % 2, R - 1, 2):
        out[k] = cur[rep_at[k]] == k + 1
    return out


# ------------------------------------------------------------- cluster ------

def _flood_roots(shape, bonds):
    n = int(np.prod(shape))
    d = len(shape)
    nb = O.neighbor_table(shape)
    adj = [[] for _ in range(n)]
    for s in range(n):
        for mu in range(d):
            if bonds[s * d + mu]:
                t = nb[s, 2 * mu]
                adj[s].append(t)
                adj[t].append(s)
    roots = [-1] * n
    for s in range(n):
        if roots[s] >= 0:
            continue
        stack, comp = [s], [s]
        roots[s] = s
        while stack:
            x = stack.pop()
            for y in adj[x]:
                if roots[y] < 0:
                    roots[y] = s
                    stack.append(y)
                    comp.append(y)
        m = min(comp)
        for x in comp:
            roots[x] = m
    return np.array(roots)


@pytest.mark.parametrize("shape", [[6], [4, 6], [3, 4, 5], [2, 3, 2, 3], [1, 5], [2, 2]])
def test_components_match_flood_fill(shape):
    # cluster.rs:323-397 (union-find vs flood fill, extent-1 and -2 wraps)
    rng = np.random.default_rng(len(shape))
    n = int(np.prod(shape))
    for frac in (0.2, 0.5, 0.8):
        bonds = (rng.random(n * len(shape)) < frac).astype(np.uint8)
        assert np.array_equal(O.components(shape, bonds), _flood_roots(shape, bonds))


def test_axis0_bonds_leave_one_cluster_per_row():
    # cluster.rs:399-415
    shape = [4, 4]
    bonds = np.zeros(32, np.uint8)
    bonds[0::2] = 1
    roots = O.components(shape, bonds)
    assert list(roots) == [0] * 4 + [4] * 4 + [8] * 4 + [12] * 4


def test_sw_limits():
    # gpu_cluster.rs:676-739: beta = 0 redraws every site on its own; a huge
    # beta labels a cold lattice as one cluster that lands on one label
    shape = [16, 16]
    d0 = O.cluster_desc(shape, q=3, beta=0.0, gap=1.0, seed=99)
    st, roots = O.cluster_sweeps(d0, np.zeros(256, np.uint8), 0, 1)
    assert np.array_equal(roots, np.arange(256))
    frac = np.bincount(st, minlength=3) / 256
    assert np.all((frac > 0.22) & (frac < 0.45))
    d8 = O.cluster_desc(shape, q=3, beta=8.0, gap=1.0, seed=99)
    st, roots = O.cluster_sweeps(d8, np.zeros(256, np.uint8), 0, 4)
    assert np.all(roots == 0) and np.all(st == st[0])


def test_wolff_limits():
    # gpu_cluster.rs:941-985: beta = 0 moves exactly one site; beta = 8
    # repaints the whole cold lattice onto a different label
    shape = [8, 8]
    d0 = O.cluster_desc(shape, q=3, seeded=True, forced=True, beta=0.0, gap=1.0, seed=5)
    st, _ = O.cluster_sweeps(d0, np.zeros(64, np.uint8), 0, 1)
    assert (st != 0).sum() == 1
    d8 = O.cluster_desc(shape, q=3, seeded=True, forced=True, beta=8.0, gap=1.0, seed=5)
    st, _ = O.cluster_sweeps(d8, np.zeros(64, np.uint8), 0, 1)
    assert np.all(st == st[0]) and st[0] != 0


def test_potts_duality_energy_at_criticality():
    # potts_e2e.rs:945-1016: at beta_c = ln(1 + sqrt(q)) the q-Potts energy
    # per site is E/N = -(1 + 1/sqrt(q)) in the thermodynamic limit; SW at
    # L = 32 lands within finite-size corrections of a few percent
    q = 3
    shape = [32, 32]
    beta = math.log(1 + math.sqrt(q))
    d = O.cluster_desc(shape, q=q, beta=beta, gap=1.0, seed=4242)
    rng = np.random.default_rng(0)
    st = rng.integers(0, q, 1024).astype(np.uint8)
    es = []
    for t in range(1500):
        st, _ = O.cluster_sweeps(d, st, t, 1)
        if t >= 100:
            agree, _ = O.potts_sums(shape, q, st)
            es.append(-agree / 1024.0)
    est = O.reduce(np.array(es))
    assert abs(est.mean - (-(1 + 1 / math.sqrt(q)))) < 0.05


# ------------------------------------------------------------ PCG64 stream -----

def test_pcg64_step_and_output_match_numpy():
    # numpy's PCG64 is the same 128-bit LCG (multiplier
    # 0x2360ed051fc65da44385df649fcccf645) with the XSL-RR output, stepped
    # before each output: from the same (state, increment) both streams agree.
    # This pins the oracle's Lcg128Xsl64 step and output; rand_core's
    # seed_from_u64 expansion and rand's f64 conversion stay unpinned.
    for seed in (0, 1, 20260728, 2**63 + 5):
        g = O.pcg64(seed)
        state = g.state_lo | (g.state_hi << 64)
        inc = g.inc_lo | (g.inc_hi << 64)
        bg = np.random.PCG64()
        bg.state = {"bit_generator": "PCG64", "state": {"state": state, "inc": inc}, "has_uint32": 0,
                    "uinteger": 0}
        want = bg.random_raw(64)
        got = [O.lib().orc_pcg64_next_u64(C.byref(g)) for _ in range(64)]
        assert [int(x) for x in want] == got
    assert O.lib().orc_pcg64_next_f64(C.byref(O.pcg64(3))) < 1.0


@pytest.mark.parametrize("shape,model,q,kernel", [
    ([16, 12], O.MODEL_ISING, 2, O.KERNEL_METROPOLIS),
    ([8, 6, 10], O.MODEL_ISING, 2, O.KERNEL_HEAT_BATH),
    ([10, 8], O.MODEL_POTTS, 4, O.KERNEL_METROPOLIS),
    ([6, 4, 4, 6], O.MODEL_Z2, 2, O.KERNEL_HEAT_BATH),
    ([8, 8, 8], O.MODEL_Z2, 2, O.KERNEL_METROPOLIS),
])
def test_threaded_sweeps_equal_the_sequential_ones(shape, model, q, kernel):
    # The all-core CPU baseline: colour passes split over threads give the
    # sequential sweep's bits, since a pass reads only the other colour.
    d = O.local_desc(shape, model=model, q=q, kernel=kernel, beta=0.7, j=1.0, h=0.2, seed=99)
    nv = int(np.prod(shape)) * (len(shape) if model == O.MODEL_Z2 else 1)
    st = np.random.default_rng(4).integers(0, q, nv).astype(np.uint8)
    want = O.local_sweeps(d, st, 3, 5)
    for threads in (1, 3, 8):
        got, secs = O.local_sweeps_mt(d, st, 3, 5, threads)
        assert np.array_equal(got, want), threads
        assert secs >= 0.0
