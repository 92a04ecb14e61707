"""Bit-exact parity of the HIP path against the CPU oracle (GPU).

Every test calls the product through the C ABI (plaquette_amd ->
libplaquette_hip.so) and compares with the oracle on the same seeded inputs:
states must be identical byte for byte, integer observables equal, cluster
partitions identical after canonical (minimum-site) relabelling.
"""
import math
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

M, HB = 0, 1


def start_config(n, q, hot, seed):
    if not hot:
        return np.zeros(n, np.uint8)
    return np.random.default_rng(seed).integers(0, q, n).astype(np.uint8)


def run_oracle_local(shape, model, q, kernel, profile, beta, j, h, offsets, seed, start, sweeps):
    d = O.local_desc(shape, model=model, q=q, kernel=kernel, profile=profile, beta=beta, j=j, h=h,
                     offsets=offsets, seed=seed)
    return O.local_sweeps(d, start, 0, sweeps)


# -- site fields: Ising ---------------------------------------------------------

ISING_CASES = [
    # shape, kernel, profile, beta, j, h, hot, sweeps
    ([32, 32, 32], M, 0, 0.2216546, 1.0, 0.0, False, 100),
    ([32, 32, 32], M, 0, 0.2216546, 1.0, 0.0, True, 100),
    ([32, 32, 32], HB, 0, 0.2216546, 1.0, 0.0, False, 100),
    ([32, 32, 32], HB, 0, 0.2216546, 1.0, 0.0, True, 100),
    ([32, 32, 32], M, 1, 0.2216546, 1.0, 0.0, True, 60),
    ([32, 32, 32], HB, 1, 0.2216546, 1.0, 0.0, True, 60),
    ([64, 64], M, 0, 0.4406868, 1.0, 0.0, True, 200),
    ([64, 64], HB, 1, 0.4406868, 1.0, 0.0, True, 200),
    ([6, 10], M, 0, 0.4406868, 1.0, 0.0, True, 200),
    ([4, 6, 8], HB, 0, 0.25, 1.0, 0.1, True, 150),
    ([4, 6, 8], M, 1, 0.25, 0.7, -0.3, True, 150),
    ([8, 4, 4, 2], M, 0, 0.15, 1.0, 0.0, True, 100),
    ([6, 4, 2, 4], HB, 0, 0.15, 1.0, 0.2, True, 100),
    ([16], M, 0, 0.8, 1.0, 0.1, True, 300),
    ([10], HB, 1, 0.8, 1.0, 0.0, True, 300),
    ([16, 8], M, 0, 0.0, 1.0, 0.0, True, 5),
    ([16, 8], HB, 0, 0.0, 1.0, 0.0, True, 5),
    ([8, 8, 8], M, 0, 5.0, 1.0, 0.0, True, 40),
]


def _id(case):
    shape, kernel, profile, beta, j, h, hot, sweeps = case
    return f"{'x'.join(map(str, shape))}-{'hb' if kernel else 'metro'}-{'f32' if profile else 'f64'}-b{beta}-{'hot' if hot else 'cold'}"


@pytest.mark.parametrize("case", ISING_CASES, ids=[_id(c) for c in ISING_CASES])
def test_ising_checkerboard_is_bit_exact(gpu, case):
    shape, kernel, profile, beta, j, h, hot, sweeps = case
    n = int(np.prod(shape))
    start = start_config(n, 2, hot, 17)
    seed = 20260728
    chain = gpu.ising_gpu_chain(shape, j, h, beta, seed, start, sweeps_between=1, kernel=kernel, profile=profile)
    # the first sweep alone, then the rest
    got1 = next(chain)
    want1 = run_oracle_local(shape, O.MODEL_ISING, 2, kernel, profile, beta, j, h, None, seed, start, 1)
    assert np.array_equal(got1, want1)
    chain.advance(sweeps - 1)
    got = chain.read()
    want = run_oracle_local(shape, O.MODEL_ISING, 2, kernel, profile, beta, j, h, None, seed, start, sweeps)
    assert chain.sweeps_done == sweeps
    assert np.array_equal(got, want), f"{(got != want).sum()} sites differ"
    # device-reduced observables of the final state equal the oracle's sums
    obs = chain.measure()
    b, s = O.ising_sums(shape, want)
    assert obs.bond_sum == b and obs.spin_sum == s
    assert obs.energy == -j * b - h * s
    chain.close()


def test_upload_readback_identity(gpu):
    # ising/gpu.rs:116-140, gpu_cluster.rs:651-664: zero sweeps is the identity
    for shape in ([4, 4], [32, 32, 32], [6, 10]):
        n = int(np.prod(shape))
        start = start_config(n, 2, True, 1)
        c = gpu.ising_gpu_chain(shape, 1.0, 0.0, 0.5, 7, start, sweeps_between=0)
        assert np.array_equal(next(c), start)
        c.close()
    st = start_config(4 * 4 * 4 * 4 * 4, 2, True, 2)
    c = gpu.gauge_gpu_chain([4, 4, 4, 4], 1.0, 0.5, 7, st, sweeps_between=0)
    assert np.array_equal(next(c), st)
    c.close()
    st = start_config(35 * 3, 3, True, 3)
    c = gpu.GpuClusterChain.swendsen_wang([5, 7, 3], 3, 1.0, 0.5, 7, st, sweeps_between=0)
    assert np.array_equal(next(c), st)
    c.close()


def test_zero_coupling_negative_field_flips_every_site_once(gpu):
    # ising/gpu.rs:142-192 at 2048^2 (4,194,304 sites): RNG-independent
    n = 2048 * 2048
    c = gpu.ising_gpu_chain([2048, 2048], 0.0, -1.0, 1.0, 7, np.zeros(n, np.uint8), sweeps_between=1)
    after = next(c)
    assert int(after.sum()) == n
    assert c.measure().spin_sum == -n
    c.close()


@pytest.mark.parametrize("kernel", [M, HB])
def test_launch_shape_does_not_change_the_bits_at_256_cubed(gpu, monkeypatch, kernel):
    # Full-size property: the 16-, 4- and 1-site-per-lane kernels are
    # independent implementations with different grids; with keyed draws they
    # must agree bit for bit at 256^3, and so must their fused observables.
    shape = [256, 256, 256]
    n = int(np.prod(shape))
    start = start_config(n, 2, True, 5)
    outs = []
    for vec in ("16", "16-rowwise", "4", "1"):
        # "16" is the uniform-parity pass at this shape; "16-rowwise" the
        # 16-site pass whose waves mix rows of both parities
        monkeypatch.setenv("PLQ_SITES_PER_LANE", vec[:2].rstrip("-"))
        monkeypatch.setenv("PLQ_X16U", "0" if vec == "16-rowwise" else "1")
        c = gpu.ising_gpu_chain(shape, 1.0, 0.0, 0.2216546, 99, start, kernel=kernel)
        c.record(4, 8)
        c.advance(12)
        outs.append((c.read(), [(o.bond_sum, o.spin_sum) for o in c.series()], c.measure().bond_sum))
        c.close()
    for other in outs[1:]:
        assert np.array_equal(outs[0][0], other[0])
        assert outs[0][1] == other[1] and outs[0][2] == other[2]
    assert outs[0][1][-1][0] == outs[0][2]


@pytest.mark.parametrize("vec", ["16", "4", "1"])
def test_every_launch_shape_is_bit_exact_against_the_oracle(gpu, monkeypatch, vec):
    monkeypatch.setenv("PLQ_SITES_PER_LANE", vec)
    for shape, kernel in (([64, 32, 32], M), ([32, 64], HB), ([32, 8, 16, 8], M)):
        n = int(np.prod(shape))
        start = start_config(n, 2, True, 12)
        c = gpu.ising_gpu_chain(shape, 1.0, 0.05, 0.3, 5, start, kernel=kernel)
        c.record(1, 30)
        c.advance(30)
        got = c.read()
        series = c.series()
        c.close()
        d = O.local_desc(shape, kernel=kernel, beta=0.3, h=0.05, seed=5)
        st = start
        for t in range(30):
            st = O.local_sweeps(d, st, t, 1)
            assert (series[t].bond_sum, series[t].spin_sum) == O.ising_sums(shape, st)
        assert np.array_equal(got, st)


# Shapes the uniform-parity 16-site pass serves (L1 >= 2 * 64 / (L0 / 32)),
# and one whose f32 thresholds underflow to 0 at beta = 5 (that chain falls
# back to the row-wise 16-site pass)
X16U_CASES = [
    # shape, kernel, profile, beta, h
    ([256, 16, 8], M, 0, 0.2216546, 0.07),
    ([256, 16, 8], HB, 1, 0.2216546, -0.05),
    ([64, 64, 4], HB, 0, 0.3, 0.0),
    ([128, 32, 4, 2], M, 1, 0.15, 0.1),
    ([128, 32, 4, 2], HB, 0, 0.2216546, 0.05),  # 4D: 9 neighbour classes per spin state
    ([64, 64, 2, 2], HB, 1, 0.3, 0.0),
    ([64, 64, 2, 2], M, 0, 0.1, 0.3),          # strong field: the 8-aligned classes differ per spin
    ([512, 32], HB, 1, 0.4406868, 0.0),
    ([2048, 8], M, 0, 0.4406868, 0.2),
    ([256, 16, 8], M, 1, 5.0, 0.0),
]


@pytest.mark.parametrize("case", X16U_CASES, ids=[f"{'x'.join(map(str, c[0]))}-{c[1]}-{c[2]}-b{c[3]}" for c in X16U_CASES])
def test_uniform_parity_pass_is_bit_exact(gpu, monkeypatch, case):
    shape, kernel, profile, beta, h = case
    n = int(np.prod(shape))
    for hot in (True, False):
        start = start_config(n, 2, hot, 14)
        c = gpu.ising_gpu_chain(shape, 1.0, h, beta, 8, start, kernel=kernel, profile=profile)
        c.record(1, 25)
        c.advance(25)
        got, series = c.read(), c.series()
        c.close()
        d = O.local_desc(shape, kernel=kernel, profile=profile, beta=beta, h=h, seed=8)
        st = start
        for t in range(25):
            st = O.local_sweeps(d, st, t, 1)
            assert (series[t].bond_sum, series[t].spin_sum) == O.ising_sums(shape, st), t
        assert np.array_equal(got, st), f"{(got != st).sum()} sites differ"


GRAPH_CASES = [
    # shape, kernel, record every (0: off), advances
    ([256, 16, 8], M, 10, [3, 7, 25, 40, 1, 24]),
    ([512, 32], HB, 1, [5, 2, 9]),
    ([64, 64, 4], M, 0, [15, 33, 16, 2]),
    ([128, 32, 4, 2], HB, 64, [70, 64, 130]),
    ([256, 16, 8], HB, 100, [150, 60]),  # beyond 64 sweeps per record: no replays
]


@pytest.mark.parametrize("case", GRAPH_CASES, ids=[f"{'x'.join(map(str, c[0]))}-{c[1]}-every{c[2]}" for c in GRAPH_CASES])
def test_graph_replays_of_site_sweeps_are_bit_exact(gpu, monkeypatch, case):
    # With PLQ_SITE_GRAPH=1, uniform-pass chains replay whole blocks of
    # sweeps (the record interval, or 16 without records) from a captured
    # graph, keyed by device counters, and launch the sweeps around them
    # directly.  Advances of odd lengths shift the phase; a set_sweep and a
    # new record ring midway invalidate the counters and the graph.  States
    # and records equal the direct launches (the default) and the oracle.
    shape, kernel, every, steps = case
    n = int(np.prod(shape))
    start = start_config(n, 2, True, 31)
    outs = []
    for graph in ("1", "0"):
        monkeypatch.setenv("PLQ_SITE_GRAPH", graph)
        c = gpu.ising_gpu_chain(shape, 1.0, 0.05, 0.2216546, 12, start, kernel=kernel)
        if every:
            c.record(every, 400)
        for k in steps:
            c.advance(k)
        mid = c.read()
        c.set_sweep(1000)
        if every:
            c.record(every, 400)
        c.advance(2 * (every or 16) + 3)
        outs.append((mid, c.read(), [(o.sweep, o.bond_sum, o.spin_sum) for o in c.series()]))
        c.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    assert outs[0][2] == outs[1][2]
    d = O.local_desc(shape, kernel=kernel, beta=0.2216546, h=0.05, seed=12)
    mid = O.local_sweeps(d, start, 0, sum(steps))
    assert np.array_equal(outs[0][0], mid)
    more = 2 * (every or 16) + 3
    assert np.array_equal(outs[0][1], O.local_sweeps(d, mid, 1000, more))
    if every:
        sweeps = [r[0] for r in outs[0][2]]
        assert sweeps == [s for s in range(1001, 1001 + more) if s % every == 0]
        for sw, bond, spin in outs[0][2]:
            assert (bond, spin) == O.ising_sums(shape, O.local_sweeps(d, mid, 1000, sw - 1000)), sw


SMALL_CASES = [([64, 64], M), ([64, 64], HB), ([16], M), ([2, 4], HB), ([16, 16, 16], M), ([8, 8, 4, 4], HB),
               ([128, 64], M), ([4, 2, 8], HB)]


@pytest.mark.parametrize("shape,kernel", SMALL_CASES, ids=[f"{'x'.join(map(str, s))}-{k}" for s, k in SMALL_CASES])
def test_small_chain_matches_per_pass_kernels(gpu, monkeypatch, shape, kernel):
    # Small power-of-two lattices run whole sweeps in one workgroup; the
    # per-pass kernels are an independent implementation.  Advances of 5,
    # 4100 and 7 sweeps cross the 4096-sweep launch limit and shift the
    # recording phase; the ring holds the last 40 of 1370 records.
    n = int(np.prod(shape))
    start = start_config(n, 2, True, 21)
    outs = []
    for small in ("1", "0"):
        monkeypatch.setenv("PLQ_SMALL_CHAIN", small)
        c = gpu.ising_gpu_chain(shape, 1.0, 0.1, 0.42, 11, start, kernel=kernel)
        c.record(3, 40)
        for k in (5, 4100, 7):
            c.advance(k)
        outs.append((c.read(), [(o.sweep, o.bond_sum, o.spin_sum) for o in c.series()], c.sweeps_done))
        c.close()
    assert outs[0][2] == outs[1][2] == 4112
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1]
    assert [s for s, _, _ in outs[0][1]] == list(range(4113 - 3 * 40, 4113, 3))


SMALL_CLUSTER_CASES = [([64, 64], 3, "sw"), ([7, 5], 2, "wolff"), ([16], 2, "sw"), ([9, 9, 9], 4, "wolff"),
                       ([4, 6, 4, 8], 3, "sw"), ([90, 91], 2, "sw"), ([32, 32], 5, "wolff")]


@pytest.mark.parametrize("shape,q,kind", SMALL_CLUSTER_CASES,
                         ids=[f"{k}-q{q}-{'x'.join(map(str, s))}" for s, q, k in SMALL_CLUSTER_CASES])
def test_small_cluster_chain_matches_the_four_kernel_sweep(gpu, monkeypatch, shape, q, kind):
    # Lattices of up to 8,192 sites run whole cluster sweeps in one
    # workgroup; the four-kernel sweep is an independent implementation.
    # Advances of 5, 4100 and 7 sweeps cross the 4096-sweep launch limit, the
    # ring keeps the last 40 of 1370 records, and the last sweep's partition
    # (roots, size histogram, relabelled sites) must agree as well.
    n = int(np.prod(shape))
    start = start_config(n, q, True, 31)
    beta = 1.0 if kind == "sw" else 0.6
    outs = []
    for small in ("1", "0"):
        monkeypatch.setenv("PLQ_SMALL_CLUSTER", small)
        make = gpu.GpuClusterChain.swendsen_wang if kind == "sw" else gpu.GpuClusterChain.wolff
        c = make(shape, q, 1.0, beta, 77, start)
        c.record(3, 40)
        for k in (5, 4100, 7):
            c.advance(k)
        st = c.stats()
        outs.append((c.read(), [(o.sweep, o.bond_sum, tuple(o.counts[a] for a in range(q))) for o in c.series()],
                     c.roots(), (st.moved, st.n_clusters, st.largest, tuple(st.size_hist)), c.sweeps_done))
        c.close()
    assert outs[0][4] == outs[1][4] == 4112
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1]
    assert np.array_equal(outs[0][2], outs[1][2])
    assert outs[0][3] == outs[1][3]
    assert [s for s, _, _ in outs[0][1]] == list(range(4113 - 3 * 40, 4113, 3))


def test_recorded_series_matches_per_sweep_measurement(gpu):
    shape = [64, 64]
    n = 4096
    start = start_config(n, 2, True, 8)
    c = gpu.ising_gpu_chain(shape, 1.0, 0.0, 0.4406868, 3, start)
    c.record(1, 50)
    c.advance(30)
    series = c.series()
    assert [o.sweep for o in series] == list(range(1, 31))
    st = start.copy()
    d = O.local_desc(shape, beta=0.4406868, seed=3)
    for t, o in enumerate(series):
        st = O.local_sweeps(d, st, t, 1)
        b, s = O.ising_sums(shape, st)
        assert (o.bond_sum, o.spin_sum) == (b, s)
    c.close()


# -- site fields: Potts ------------------------------------------------------------

POTTS_CASES = [
    ([8, 8], 3, M, 0, 1.0050525, None),
    ([8, 8], 3, HB, 0, 1.0050525, None),
    ([6, 4, 4], 4, M, 1, 0.6, [0.0, 0.1, -0.2, 0.05]),
    ([6, 4, 4], 3, HB, 1, 0.6, [0.3, 0.0, -0.1]),
    ([10, 6], 5, M, 0, 0.9, None),
]


@pytest.mark.parametrize("case", POTTS_CASES)
def test_potts_checkerboard_is_bit_exact(gpu, case):
    shape, q, kernel, profile, beta, offsets = case
    n = int(np.prod(shape))
    start = start_config(n, q, True, 4)
    c = gpu.potts_gpu_chain(shape, q, 1.0, offsets, beta, 11, start, kernel=kernel, profile=profile)
    c.advance(80)
    got = c.read()
    want = run_oracle_local(shape, O.MODEL_POTTS, q, kernel, profile, beta, 1.0, 0.0, offsets, 11, start, 80)
    assert np.array_equal(got, want)
    agree, counts = O.potts_sums(shape, q, want)
    obs = c.measure()
    assert obs.bond_sum == agree and list(obs.counts[:q]) == list(counts)
    c.close()


# -- link fields: Z2 -------------------------------------------------------------

Z2_CASES = [
    ([8, 8], M, 0, 0.5),
    ([8, 8], HB, 0, 0.5),
    ([4, 6, 4], M, 0, 0.75),
    ([4, 6, 4], HB, 1, 0.75),
    ([4, 4, 6, 4], HB, 0, 0.44),
    ([4, 4, 6, 4], M, 1, 0.44),
    ([8, 8, 8, 8], HB, 0, 0.44),
]


@pytest.mark.parametrize("case", Z2_CASES)
def test_z2_link_checkerboard_is_bit_exact(gpu, case):
    shape, kernel, profile, beta = case
    n = int(np.prod(shape)) * len(shape)
    start = start_config(n, 2, True, 6)
    c = gpu.gauge_gpu_chain(shape, 1.0, beta, 21, start, kernel=kernel, profile=profile)
    c.advance(40)
    got = c.read()
    want = run_oracle_local(shape, O.MODEL_Z2, 2, kernel, profile, beta, 1.0, 0.0, None, 21, start, 40)
    assert np.array_equal(got, want)
    assert c.measure().bond_sum == O.plaquette_sum(shape, want)
    c.close()


# -- cluster ---------------------------------------------------------------------

CLUSTER_CASES = [
    # shape, q, gap, beta, seeded, forced, profile, sweeps
    ([64, 64], 3, 1.0, math.log(1 + math.sqrt(3)), False, False, 0, 20),
    ([64, 64], 3, 1.0, math.log(1 + math.sqrt(3)), False, False, 1, 10),
    ([300, 200], 3, 1.0, 1.0050525, False, False, 0, 5),
    ([5, 7, 3], 3, 1.0, 0.5, False, False, 0, 20),
    ([16, 16, 16], 2, 2.0, 0.2216546, False, False, 0, 10),
    ([6, 5, 4, 3], 4, 1.0, 0.7, False, True, 0, 10),
    ([32, 32], 3, 1.0, 1.2, True, True, 0, 30),
    ([32, 32], 2, 2.0, 0.4406868, True, True, 1, 30),
    ([1000], 3, 1.0, 2.0, False, False, 0, 10),
    ([256, 256], 3, 1.0, 1.0050525, False, False, 0, 4),   # power-of-two boxes, 4 of 128x128
    ([32, 32, 32], 2, 2.0, 0.2216546, True, True, 0, 6),   # power-of-two boxes, 2 of 32x32x16
    ([16, 2048], 2, 2.0, 0.4406868, False, False, 0, 4),  # 16x1024 boxes: one segment per row, axis-0 wrap
    ([512, 64], 3, 1.0, 1.0050525, True, True, 0, 6),     # 256x64 boxes spanning axis 1
    ([16, 8, 8, 16], 3, 1.0, 0.9, False, False, 0, 5),    # one box spanning every axis
    ([65536], 2, 2.0, 1.5, False, False, 0, 4),           # four 1D boxes
    ([256, 256], 4, 1.0, 0.01, False, False, 0, 2),       # nearly all singletons: the size table overflows
]


@pytest.mark.parametrize("case", CLUSTER_CASES)
def test_cluster_sweeps_are_bit_exact(gpu, case):
    shape, q, gap, beta, seeded, forced, profile, sweeps = case
    n = int(np.prod(shape))
    start = start_config(n, q, True, 9)
    extent = gpu.Extent.SEEDED if seeded else gpu.Extent.ALL
    relabel = gpu.Relabel.FORCED_CHANGE if forced else gpu.Relabel.REDRAW
    c = gpu.GpuClusterChain.new(shape, q, gap, extent, relabel, beta, 31, start, profile=profile)
    d = O.cluster_desc(shape, q=q, seeded=seeded, forced=forced, profile=profile, beta=beta, gap=gap, seed=31)
    st = start
    for t in range(sweeps):
        c.advance(1)
        st, roots = O.cluster_sweeps(d, st, t, 1)
        got = c.read()
        assert np.array_equal(got, st), f"sweep {t}: {(got != st).sum()} sites differ"
        if t in (0, sweeps - 1):
            assert np.array_equal(c.roots(), roots)
    stats = c.stats()
    sizes = np.bincount(roots)
    sizes = sizes[sizes > 0]
    assert stats.n_clusters == sizes.size and stats.largest == sizes.max()
    hist = np.bincount(np.floor(np.log2(sizes)).astype(int), minlength=40)
    assert list(stats.size_hist) == list(hist)
    agree, counts = O.potts_sums(shape, q, st)
    obs = c.measure()
    assert obs.bond_sum == agree and list(obs.counts[:q]) == list(counts)
    c.close()


MULTI_SWEEP_CASES = [c for c in CLUSTER_CASES if all(e & (e - 1) == 0 for e in c[0]) and np.prod(c[0]) > 8192]


@pytest.mark.parametrize("case", MULTI_SWEEP_CASES)
def test_cluster_advance_of_many_sweeps_is_bit_exact(gpu, case):
    # One advance of several sweeps fuses each sweep's relabel with the next
    # sweep's box labelling; every recorded sweep must still match the
    # oracle's sweep by sweep, and the partition of the last sweep its roots.
    shape, q, gap, beta, seeded, forced, profile, sweeps = case
    n = int(np.prod(shape))
    start = start_config(n, q, True, 13)
    extent = gpu.Extent.SEEDED if seeded else gpu.Extent.ALL
    relabel = gpu.Relabel.FORCED_CHANGE if forced else gpu.Relabel.REDRAW
    c = gpu.GpuClusterChain.new(shape, q, gap, extent, relabel, beta, 37, start, profile=profile)
    d = O.cluster_desc(shape, q=q, seeded=seeded, forced=forced, profile=profile, beta=beta, gap=gap, seed=37)
    k = sweeps + 3
    c.record(1, k)
    c.advance(k)
    series = c.series()
    st = start
    for t in range(k):
        st, roots = O.cluster_sweeps(d, st, t, 1)
        agree, counts = O.potts_sums(shape, q, st)
        assert series[t].bond_sum == agree and list(series[t].counts[:q]) == list(counts), f"sweep {t}"
    assert np.array_equal(c.read(), st)
    assert np.array_equal(c.roots(), roots)
    c.close()


RELABEL_MEASURE_CASES = [
    # shape, q, beta, sweeps
    ([256, 256], 3, 1.0050525, 6),    # 4 boxes of 128x128: halo rows and columns
    ([32, 32, 32], 2, 0.2216546, 5),  # boxes of 32x32x16: the halo of axis 2 only
    ([16, 2048], 2, 0.4406868, 5),    # boxes as wide as the lattice: axis 0 wraps inside the box
    ([65536], 3, 2.0, 5),             # four 1D boxes: one halo byte each
    ([16, 8, 8, 16], 3, 0.9, 5),      # one box spanning every axis: no halo
    ([512, 256], 16, 1.2, 4),         # q = 16, every count column
]


@pytest.mark.parametrize("shape,q,beta,sweeps", RELABEL_MEASURE_CASES,
                         ids=["x".join(map(str, c[0])) + f"-q{c[1]}" for c in RELABEL_MEASURE_CASES])
def test_sw_record_counted_in_the_relabel(gpu, monkeypatch, shape, q, beta, sweeps):
    # A recorded Swendsen-Wang sweep counts its record in the per-box
    # relabel, reading the new labels of forward neighbours in other boxes
    # from their cluster roots; PLQ_RELABEL_MEASURE=0 records with the
    # separate measurement pass.  Both series, sweep by sweep, against the
    # oracle's sums, and the labels against each other.
    n = int(np.prod(shape))
    start = start_config(n, q, True, 17)
    outs = []
    for fused in ("1", "0"):
        monkeypatch.setenv("PLQ_RELABEL_MEASURE", fused)
        c = gpu.GpuClusterChain.swendsen_wang(shape, q, 1.0, beta, 23, start)
        c.record(1, sweeps)
        c.advance(sweeps)
        outs.append(([(o.sweep, o.bond_sum, tuple(o.counts[a] for a in range(q))) for o in c.series()], c.read()))
        c.close()
    assert outs[0][0] == outs[1][0]
    assert np.array_equal(outs[0][1], outs[1][1])
    d = O.cluster_desc(shape, q=q, beta=beta, gap=1.0, seed=23)
    st = start
    for t in range(sweeps):
        st, _ = O.cluster_sweeps(d, st, t, 1)
        agree, counts = O.potts_sums(shape, q, st)
        assert outs[0][0][t][1:] == (agree, tuple(counts)), f"sweep {t}"
    assert np.array_equal(outs[0][1], st)


@pytest.mark.parametrize("shape,q,seeded", [([256, 256], 3, False), ([32, 32, 32], 2, True), ([16, 2048], 2, False),
                                            ([90, 91], 3, False)])
def test_cluster_stats_after_every_sweep(gpu, shape, q, seeded):
    # The size histogram taken after every sweep (box-counted where the box
    # kernels apply, with the cross-box counts cleared between calls) equals
    # the oracle's partition of that sweep.
    n = int(np.prod(shape))
    beta = 1.0050525 if q == 3 else 0.4406868
    start = start_config(n, q, True, 21)
    extent = gpu.Extent.SEEDED if seeded else gpu.Extent.ALL
    relabel = gpu.Relabel.FORCED_CHANGE if seeded else gpu.Relabel.REDRAW
    c = gpu.GpuClusterChain.new(shape, q, 1.0, extent, relabel, beta, 41, start)
    d = O.cluster_desc(shape, q=q, seeded=seeded, forced=seeded, beta=beta, gap=1.0, seed=41)
    st = start
    for t in range(4):
        c.advance(1)
        st, roots = O.cluster_sweeps(d, st, t, 1)
        for _ in range(2):  # twice: the second call sees the cleared counts
            stats = c.stats()
            sizes = np.bincount(roots)
            sizes = sizes[sizes > 0]
            assert stats.n_clusters == sizes.size and stats.largest == sizes.max(), t
            hist = np.bincount(np.floor(np.log2(sizes)).astype(int), minlength=40)
            assert list(stats.size_hist) == list(hist), t
    c.close()


@pytest.mark.parametrize("shape,q,seeded", [([256, 256], 3, False), ([32, 32, 32], 2, True), ([90, 91], 3, False),
                                            ([64, 64], 3, False)])
def test_recorded_cluster_statistics_match_every_sweep(gpu, shape, q, seeded):
    # Cluster statistics kept on the device for every recorded sweep of one
    # multi-sweep advance (the small-lattice chain included: it then runs the
    # four-kernel sweep) equal the oracle's partition of each sweep.
    n = int(np.prod(shape))
    beta = 1.0050525 if q == 3 else 0.4406868
    start = start_config(n, q, True, 23)
    extent = gpu.Extent.SEEDED if seeded else gpu.Extent.ALL
    relabel = gpu.Relabel.FORCED_CHANGE if seeded else gpu.Relabel.REDRAW
    c = gpu.GpuClusterChain.new(shape, q, 1.0, extent, relabel, beta, 43, start)
    d = O.cluster_desc(shape, q=q, seeded=seeded, forced=seeded, beta=beta, gap=1.0, seed=43)
    k = 6
    c.advance(2)
    st, _ = O.cluster_sweeps(d, start, 0, 2)
    c.record(1, k)
    c.record_clusters(True)
    c.advance(k)
    got = c.cluster_series()
    assert len(got) == k
    for t in range(k):
        st, roots = O.cluster_sweeps(d, st, 2 + t, 1)
        sizes = np.bincount(roots)
        sizes = sizes[sizes > 0]
        s = got[t]
        assert s.sweep == 3 + t
        assert s.n_clusters == sizes.size and s.largest == sizes.max(), t
        assert list(s.size_hist) == list(np.bincount(np.floor(np.log2(sizes)).astype(int), minlength=40)), t
        if not seeded:
            assert s.moved == n
    c.close()


@pytest.mark.parametrize("shape,q,seeded", [([256, 256], 3, False), ([32, 32, 32], 2, True), ([65536], 2, False),
                                            ([16, 8, 8, 16], 3, False)])
def test_recorded_cluster_statistics_every_third_sweep_with_calls_between(gpu, shape, q, seeded):
    # A record every third sweep: the recorded sweeps count their clusters
    # in the relabel, the others do not.  A statistics call after every
    # sweep, recorded or not, and the kept series both equal the oracle.
    n = int(np.prod(shape))
    beta = 1.0050525 if q == 3 else 0.4406868
    start = start_config(n, q, True, 29)
    extent = gpu.Extent.SEEDED if seeded else gpu.Extent.ALL
    relabel = gpu.Relabel.FORCED_CHANGE if seeded else gpu.Relabel.REDRAW
    c = gpu.GpuClusterChain.new(shape, q, 1.0, extent, relabel, beta, 47, start)
    d = O.cluster_desc(shape, q=q, seeded=seeded, forced=seeded, beta=beta, gap=1.0, seed=47)
    c.record(3, 8)
    c.record_clusters(True)
    st, want = start, []

    def summary(roots):
        sizes = np.bincount(roots)
        sizes = sizes[sizes > 0]
        return sizes.size, int(sizes.max()), np.bincount(np.floor(np.log2(sizes)).astype(int), minlength=40).tolist()

    for t in range(9):
        c.advance(1)
        st, roots = O.cluster_sweeps(d, st, t, 1)
        s = c.stats()
        assert (s.n_clusters, s.largest, list(s.size_hist)) == summary(roots), t
        if (t + 1) % 3 == 0:
            want.append((t + 1, summary(roots)))
    got = c.cluster_series()
    assert [g.sweep for g in got] == [w[0] for w in want]
    for g, (_, w) in zip(got, want):
        assert (g.n_clusters, g.largest, list(g.size_hist)) == w
    c.close()


def test_cluster_budget_exhaustion_is_reported(gpu):
    # gpu_cluster.rs:512-518: labelling past its cap is an error, forced from
    # the host with a zero hop budget on an ordered (percolating) lattice
    n = 64 * 64
    c = gpu.GpuClusterChain.swendsen_wang([64, 64], 3, 1.0, 8.0, 1, np.zeros(n, np.uint8))
    c.set_pass_cap(0)
    c.advance(1)  # enqueue only: the overrun is reported where the chain is waited for
    with pytest.raises(gpu.PlaquetteError) as err:
        c.sync()
    assert err.value.code == gpu.PLQ_ENOCONV
    with pytest.raises(gpu.PlaquetteError) as err:  # and stays: the chain holds no configuration
        c.read()
    assert err.value.code == gpu.PLQ_ENOCONV
    c.close()


# -- physics: 2D Ising against Kaufman's exact finite-torus solution -------------

@pytest.mark.parametrize("L,beta,kernel", [(64, 0.4406868, M), (64, 0.4406868, HB), (32, 0.3, M), (32, 0.6, HB)])
def test_ising2d_energy_and_specific_heat_match_kaufman(gpu, L, beta, kernel):
    import exact_ising2d as X
    from plaquette_amd import statistics as S

    n = L * L
    start = start_config(n, 2, True, 77)
    c = gpu.ising_gpu_chain([L, L], 1.0, 0.0, beta, 1234, start, kernel=kernel)
    c.advance(20000)
    steps = 300000
    c.record(1, steps)
    c.advance(steps)
    e = np.array([o.energy for o in c.series()])
    c.close()
    est = S.reduce(e / n)
    want = X.energy_per_site(L, L, beta)
    assert abs(est.mean - want) < 4.5 * est.stderr, (est.mean, want, est.stderr)
    cv = S.specific_heat(e, beta, n)
    want_c = X.specific_heat(L, L, beta)
    assert abs(cv.value - want_c) < 4.5 * cv.stderr, (cv.value, want_c, cv.stderr)


# -- the CPU stream: the reference CPU path's own PCG64 draws ---------------------

PMODE_CASES = [
    # shape, kernel, beta, j, h, sweeps (in advances)
    ([16, 16], M, 0.4406868, 1.0, 0.0, (1, 4, 15)),
    ([16, 16], HB, 0.4406868, 1.0, 0.1, (1, 4, 15)),
    ([6, 10], M, 0.3, 1.0, -0.2, (2, 8)),
    ([4, 6, 8], HB, 0.25, 0.8, 0.05, (3, 7)),
    ([8, 4, 4, 2], M, 0.15, 1.0, 0.0, (5, 5)),
    ([16], M, 0.8, 1.0, 0.1, (10, 10)),
    ([64, 64, 32], M, 0.2216546, 1.0, 0.0, (1, 2)),   # 131,072 sites: several blocks per pass
    ([128, 128], HB, 0.4406868, 1.0, 0.0, (2, 2)),
]


@pytest.mark.parametrize("case", PMODE_CASES, ids=[f"{'x'.join(map(str, c[0]))}-{c[1]}" for c in PMODE_CASES])
def test_cpu_stream_profile_reproduces_the_cpu_path(gpu, case):
    # The CPU path as written (updater.rs:77-167, 316-324; RandRng = PCG64
    # seed_from_u64, rng.rs:30-50), restated sequentially by the oracle, and
    # the device's scan + jump-ahead reproduction of the same stream.
    shape, kernel, beta, j, h, advances = case
    n = int(np.prod(shape))
    start = start_config(n, 2, True, 61)
    seed = 20260728
    c = gpu.ising_gpu_chain(shape, j, h, beta, seed, start, kernel=kernel, profile=gpu.Profile.CPU_STREAM)
    g = O.pcg64(seed)
    assert c.pcg_state() == (g.state_lo | (g.state_hi << 64), g.inc_lo | (g.inc_hi << 64))
    done = 0
    for k in advances:
        c.advance(k)
        done += k
        st, _ = O.pmode_ising_sweeps(shape, kernel, beta, j, h, seed, start, done)
        got = c.read()
        assert np.array_equal(got, st), f"{(got != st).sum()} sites differ"
    c.close()


def test_cpu_stream_state_continues_in_a_new_chain(gpu):
    shape, beta = [32, 32], 0.4406868
    n = 1024
    start = start_config(n, 2, True, 62)
    a = gpu.ising_gpu_chain(shape, 1.0, 0.0, beta, 7, start, profile=gpu.Profile.CPU_STREAM)
    a.advance(5)
    mid, state = a.read(), a.pcg_state()
    a.advance(6)
    b = gpu.ising_gpu_chain(shape, 1.0, 0.0, beta, 999, mid, profile=gpu.Profile.CPU_STREAM)
    b.set_pcg_state(*state)
    b.advance(6)
    assert np.array_equal(a.read(), b.read())
    assert a.pcg_state() == b.pcg_state()
    a.close()
    b.close()


# -- the configurations at their full sizes ------------------------------------------

@pytest.mark.parametrize("start_kind,stats", [("cold", True), ("hot", True), ("hot", False)])
def test_potts_sw_at_4096_squared_is_bit_exact(gpu, start_kind, stats):
    # q = 3 Swendsen-Wang at beta_c on the full 4096^2 lattice (1,024 boxes
    # of 128x128).  From a cold start the first sweeps have one percolating
    # cluster whose root every box's boundary unions meet.  Labels, the last
    # partition, the recorded observables and (with stats) the recorded
    # cluster statistics of every sweep against the oracle.  Without the
    # statistics the relabel counts each record itself.
    shape, q, beta = [4096, 4096], 3, 1.0050525
    n = 4096 * 4096
    start = start_config(n, q, start_kind == "hot", 5)
    k = 3
    c = gpu.GpuClusterChain.new(shape, q, 1.0, gpu.Extent.ALL, gpu.Relabel.REDRAW, beta, 2026, start)
    c.record(1, k)
    c.record_clusters(stats)
    c.advance(k)
    got_series, got_stats = c.series(), (c.cluster_series() if stats else None)
    got, got_roots = c.read(), c.roots()
    c.close()
    d = O.cluster_desc(shape, q=q, beta=beta, gap=1.0, seed=2026)
    st = start
    largest = 0
    for t in range(k):
        st, roots = O.cluster_sweeps(d, st, t, 1)
        agree, counts = O.potts_sums(shape, q, st)
        assert got_series[t].bond_sum == agree and list(got_series[t].counts[:q]) == list(counts), t
        sizes = np.bincount(roots)
        sizes = sizes[sizes > 0]
        largest = max(largest, int(sizes.max()))
        if not stats:
            continue
        s = got_stats[t]
        assert s.n_clusters == sizes.size and s.largest == sizes.max(), t
        hist = np.bincount(np.floor(np.log2(sizes)).astype(int), minlength=40)
        assert list(s.size_hist) == list(hist), t
    assert np.array_equal(got, st)
    assert np.array_equal(got_roots, roots)
    if start_kind == "cold":
        assert largest > n // 2  # a percolating cluster was met


@pytest.mark.parametrize("kernel", [M, HB])
def test_ising_at_256_cubed_is_bit_exact(gpu, kernel):
    # The headline configuration: 3D Ising 256^3, beta = 0.2216546, cold
    # start, the default launch shape (sixteen sites per lane) against
    # the oracle's sequential sweeps, with the fused E and M records.
    shape = [256, 256, 256]
    n = 256 ** 3
    start = np.zeros(n, np.uint8)
    c = gpu.ising_gpu_chain(shape, 1.0, 0.0, 0.2216546, 20260728, start, kernel=kernel)
    c.record(1, 4)
    c.advance(2)
    series = c.series()
    got = c.read()
    c.close()
    d = O.local_desc(shape, kernel=kernel, beta=0.2216546, seed=20260728)
    st = start
    for t in range(2):
        st = O.local_sweeps(d, st, t, 1)
        b, m = O.ising_sums(shape, st)
        assert series[t].bond_sum == b and series[t].spin_sum == m, t
    assert np.array_equal(got, st), f"{(got != st).sum()} sites differ"


FUSED_CASES = [
    # shape, kernel, beta, record interval, advances
    ([256, 32, 32], M, 0.2216546, 3, (4, 5)),     # 4 tiles of 16 x 16 rows, every halo wraps
    ([256, 32, 32], HB, 0.2216546, 3, (4, 5)),
    ([256, 64, 32], M, 0.3, 2, (3, 4)),
    ([256, 32, 64], HB, 0.15, 4, (9,)),
    ([256, 32, 32], M, 0.5, 1, (3, 2)),          # a record every sweep: folds and partials alternate halves
    ([256, 256, 256], M, 0.2216546, 2, (4,)),     # the headline lattice: on by default (256 tiles)
    ([256, 256, 256], HB, 0.2216546, 3, (3,)),
]


@pytest.mark.parametrize("shape,kernel,beta,every,advances", FUSED_CASES,
                         ids=["x".join(map(str, c[0])) + ("-m" if c[1] == M else "-hb") for c in FUSED_CASES])
def test_fused_two_colour_sweep_is_bit_exact(gpu, monkeypatch, shape, kernel, beta, every, advances):
    # One launch per 3D sweep: colour 0 on each 16 x 16-row tile and its face
    # halos, kept in LDS, then colour 1 of the tile; a recorded sweep also
    # writes its partial sums, folded by the next sweep's launch.  Against
    # the per-pass launches (PLQ_FUSED_SWEEP=0) and the oracle's sequential
    # sweeps, states and records.
    n = int(np.prod(shape))
    start = start_config(n, 2, True, 41)
    sweeps = sum(advances)
    outs = []
    for fused in ("1", "0"):
        monkeypatch.setenv("PLQ_FUSED_SWEEP", fused)
        c = gpu.ising_gpu_chain(shape, 1.0, 0.0, beta, 19, start, kernel=kernel)
        c.record(every, sweeps)
        for k in advances:
            c.advance(k)
        outs.append((c.read(), [(o.sweep, o.bond_sum, o.spin_sum) for o in c.series()]))
        c.close()
    assert np.array_equal(outs[0][0], outs[1][0]), f"{(outs[0][0] != outs[1][0]).sum()} sites differ"
    assert outs[0][1] == outs[1][1]
    d = O.local_desc(shape, kernel=kernel, beta=beta, seed=19)
    st = start
    recs = []
    for t in range(sweeps):
        st = O.local_sweeps(d, st, t, 1)
        if (t + 1) % every == 0:
            recs.append((t + 1,) + tuple(O.ising_sums(shape, st)))
    assert outs[0][1] == recs
    assert np.array_equal(outs[0][0], st), f"{(outs[0][0] != st).sum()} sites differ from the oracle"


def test_fused_sweep_chain_state_after_writes_and_batches(gpu, monkeypatch):
    # The fused sweep swaps the chain's two pairs of colour arrays every
    # sweep; reads, writes, measurements and next() must see the current
    # pair.  Heat bath on 256x32x32, against the oracle throughout.
    shape, beta, seed = [256, 32, 32], 0.3, 23
    n = int(np.prod(shape))
    monkeypatch.setenv("PLQ_FUSED_SWEEP", "1")
    start = start_config(n, 2, True, 51)
    c = gpu.ising_gpu_chain(shape, 1.0, 0.0, beta, seed, start, kernel=HB, sweeps_between=3)
    d = O.local_desc(shape, kernel=HB, beta=beta, seed=seed)
    c.advance(1)  # an odd number of swaps
    st = O.local_sweeps(d, start, 0, 1)
    assert np.array_equal(c.read(), st)
    obs = c.measure()
    b, m = O.ising_sums(shape, st)
    assert (obs.bond_sum, obs.spin_sum) == (b, m)
    flipped = st ^ 1  # a write between sweeps replaces the state
    c.write(flipped)
    got = next(c)  # three more sweeps
    st = O.local_sweeps(d, flipped, 1, 3)
    assert np.array_equal(got, st)
    c.advance(2)
    st = O.local_sweeps(d, st, 4, 2)
    assert np.array_equal(c.read(), st)
    c.close()


def test_cluster_statistics_across_a_reused_crossing_tag(gpu):
    # The crossing marks carry a tag 1 + sweep mod 255, so sweeps t and
    # t + 255 share a tag and a stale mark from 255 sweeps earlier is seen
    # again.  Statistics of every one of 260 recorded sweeps at 256^2 (four
    # boxes) against the oracle's partitions.
    shape, q, beta = [256, 256], 3, 1.0050525
    n = 256 * 256
    start = start_config(n, q, True, 29)
    k = 260
    c = gpu.GpuClusterChain.new(shape, q, 1.0, gpu.Extent.ALL, gpu.Relabel.REDRAW, beta, 47, start)
    c.record(1, k)
    c.record_clusters(True)
    c.advance(k)
    got = c.cluster_series()
    final = c.read()
    c.close()
    assert len(got) == k
    d = O.cluster_desc(shape, q=q, beta=beta, gap=1.0, seed=47)
    st = start
    for t in range(k):
        st, roots = O.cluster_sweeps(d, st, t, 1)
        sizes = np.bincount(roots)
        sizes = sizes[sizes > 0]
        s = got[t]
        assert s.n_clusters == sizes.size and s.largest == sizes.max(), t
        hist = np.bincount(np.floor(np.log2(sizes)).astype(int), minlength=40)
        assert list(s.size_hist) == list(hist), t
    assert np.array_equal(final, st)
