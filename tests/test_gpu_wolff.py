"""Wolff moves grown from the seed site on the device (plq_wolff.hip): bit-exact
against the oracle's Wolff move, which decomposes the whole lattice and
relabels the seed's cluster (cluster_relabel.wgsl:55-66), whichever path each
sweep takes -- a grown cluster, a growth that gives up (level or queue
budget) and falls back to the full decomposition, or an adaptive skip -- and
the partition of the last sweep (roots, statistics) rebuilt on demand after
a grown sweep (GPU)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

BETA_C2 = 0.4406868  # 2D Ising, gap 2J

WOLFF_CASES = [
    # id, shape, q, gap, beta, forced, (mode, levels, queue), sweeps, expect
    ("hot2d", [128, 128], 2, 2.0, 0.20, True, (2, 64, None), 40, "grown"),
    ("critical2d", [128, 128], 2, 2.0, BETA_C2, True, (2, 12, None), 40, "both"),
    ("ordered2d", [256, 256], 2, 2.0, 0.60, True, (2, 256, 2000), 12, "full"),
    ("potts3d", [32, 32, 16], 3, 1.0, 0.45, True, (2, 32, None), 30, "any"),
    ("redraw_odd", [90, 110], 3, 1.0, 0.9, False, (2, 48, None), 30, "any"),
    ("hot4d", [10, 10, 10, 10], 2, 2.0, 0.08, True, (2, 32, None), 30, "grown"),
    ("ring1d", [20000], 2, 2.0, 1.2, True, (2, 64, None), 30, "any"),
    ("adaptive", [128, 128], 2, 2.0, BETA_C2, True, (1, 6, None), 60, "any"),
]


def _chain(gpu, monkeypatch, shape, q, gap, beta, forced, mode, seed, start):
    monkeypatch.setenv("PLQ_WOLFF_GROW", str(mode[0]))
    if mode[1] is not None:
        monkeypatch.setenv("PLQ_WOLFF_LEVELS", str(mode[1]))
    else:
        monkeypatch.delenv("PLQ_WOLFF_LEVELS", raising=False)
    if mode[2] is not None:
        monkeypatch.setenv("PLQ_WOLFF_QUEUE", str(mode[2]))
    else:
        monkeypatch.delenv("PLQ_WOLFF_QUEUE", raising=False)
    relabel = gpu.Relabel.FORCED_CHANGE if forced else gpu.Relabel.REDRAW
    return gpu.GpuClusterChain.new(shape, q, gap, gpu.Extent.SEEDED, relabel, beta, seed, start)


def _check_stats(stats, roots):
    sizes = np.bincount(roots)
    sizes = sizes[sizes > 0]
    assert stats.n_clusters == sizes.size and stats.largest == sizes.max()
    hist = np.bincount(np.floor(np.log2(sizes)).astype(int), minlength=40)
    assert list(stats.size_hist) == list(hist)


@pytest.mark.parametrize("case", WOLFF_CASES, ids=[c[0] for c in WOLFF_CASES])
def test_grown_wolff_sweeps_are_bit_exact(gpu, monkeypatch, case):
    _, shape, q, gap, beta, forced, mode, sweeps, expect = case
    n = int(np.prod(shape))
    start = np.random.default_rng(5).integers(0, q, n).astype(np.uint8)
    if expect == "full":
        start = np.zeros(n, np.uint8)  # ordered: the seed's cluster outgrows the queue
    c = _chain(gpu, monkeypatch, shape, q, gap, beta, forced, mode, 23, start)
    d = O.cluster_desc(shape, q=q, seeded=True, forced=forced, beta=beta, gap=gap, seed=23)
    st = start
    for t in range(sweeps):
        c.advance(1)
        st, roots = O.cluster_sweeps(d, st, t, 1)
        got = c.read()
        assert np.array_equal(got, st), f"sweep {t}: {(got != st).sum()} sites differ"
        if t in (0, sweeps // 2, sweeps - 1):
            assert np.array_equal(c.roots(), roots), f"partition of sweep {t}"
    stats = c.stats()
    _check_stats(stats, roots)
    grown, full = c.wolff_counts()
    assert grown + full == sweeps
    if expect == "grown":
        assert grown >= sweeps * 3 // 4, (grown, full)
    elif expect == "full":
        assert full >= sweeps // 2, (grown, full)
    elif expect == "both":
        assert grown > 0 and full > 0, (grown, full)
    c.close()


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_every_growth_mode_gives_the_same_chain(gpu, monkeypatch, mode):
    # One advance of many sweeps with E and M recorded every sweep: the
    # records, the final labels and the last partition are the oracle's in
    # every mode (0: always decompose, 1: adaptive, 2: attempt every sweep).
    shape, q, beta = [192, 96], 3, 1.0050525  # q = 3 Potts at beta_c
    n = int(np.prod(shape))
    start = np.random.default_rng(8).integers(0, q, n).astype(np.uint8)
    c = _chain(gpu, monkeypatch, shape, q, 1.0, beta, True, (mode, 8, None), 41, start)
    d = O.cluster_desc(shape, q=q, seeded=True, forced=True, beta=beta, gap=1.0, seed=41)
    k = 50
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
    _check_stats(c.stats(), roots)
    grown, full = c.wolff_counts()
    if mode == 0:
        assert (grown, full) == (0, 0)
    else:
        assert grown + full == k and grown > 0
    c.close()


def test_moved_is_the_seed_cluster_size_after_a_grown_sweep(gpu, monkeypatch):
    # The relabel partials a grown sweep writes: stats().moved is the size of
    # the seed's cluster in the sweep's partition.
    shape, q, beta = [128, 128], 2, 0.3
    n = int(np.prod(shape))
    start = np.random.default_rng(12).integers(0, q, n).astype(np.uint8)
    c = _chain(gpu, monkeypatch, shape, q, 2.0, beta, True, (2, 64, None), 9, start)
    d = O.cluster_desc(shape, q=q, seeded=True, forced=True, beta=beta, gap=2.0, seed=9)
    st = start
    for t in range(6):
        prev = st
        c.advance(1)
        st, roots = O.cluster_sweeps(d, st, t, 1)
        changed = np.flatnonzero(prev != st)
        assert changed.size > 0  # forced change at q = 2 flips the cluster
        r = roots[changed[0]]
        assert np.all(roots[changed] == r)
        assert c.stats().moved == int((roots == r).sum()) == changed.size
    grown, _ = c.wolff_counts()
    assert grown == 6
    c.close()


def test_writing_the_lattice_drops_a_grown_partition(gpu, monkeypatch):
    shape = [128, 128]
    n = int(np.prod(shape))
    start = np.random.default_rng(3).integers(0, 2, n).astype(np.uint8)
    c = _chain(gpu, monkeypatch, shape, 2, 2.0, 0.25, True, (2, 64, None), 4, start)
    c.advance(3)
    c.write(start)
    with pytest.raises(gpu.PlaquetteError):
        c.roots()
    d = O.cluster_desc(shape, q=2, seeded=True, forced=True, beta=0.25, gap=2.0, seed=4)
    c.advance(1)
    st, roots = O.cluster_sweeps(d, start, 3, 1)
    assert np.array_equal(c.read(), st)
    assert np.array_equal(c.roots(), roots)
    c.close()


def test_a_lowered_hop_budget_still_reports_non_convergence(gpu, monkeypatch):
    # Growth walks no trees, so a chain with a lowered hop budget takes the
    # full decomposition and reports PLQ_ENOCONV as before.
    shape = [128, 128]
    n = int(np.prod(shape))
    start = np.zeros(n, np.uint8)
    c = _chain(gpu, monkeypatch, shape, 2, 2.0, 8.0, True, (2, 64, None), 6, start)
    c.set_pass_cap(0)
    c.advance(1)
    with pytest.raises(gpu.PlaquetteError) as err:
        c.sync()
    assert err.value.code == gpu.PLQ_ENOCONV
    c.close()
