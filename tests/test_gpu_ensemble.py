"""Batched Z2 ensembles with replica exchange: bit-exact against the oracle
(states, beta assignments, recorded plaquette sums, exchange counts) and
independent of the launch shape, of hipGraph replay and of the cooperative
one-launch-per-block mode (GPU)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def make_case(kind):
    if kind == "mixed4d":
        shapes = [[4, 4, 4, 4]] * 3 + [[6, 4, 4, 4]] * 2 + [[4, 4, 6, 6]]
        betas = [0.44, 0.40, 0.48, 0.42, 0.46, 0.45]
        return 4, shapes, betas
    if kind == "uneven3d":
        shapes = [[4, 4, 4], [6, 6, 6], [4, 4, 4], [8, 6, 4], [6, 6, 6], [4, 4, 4]]
        betas = [0.7, 0.8, 0.75, 0.9, 0.85, 0.8]
        return 3, shapes, betas
    shapes = [[6, 6]] * 4 + [[4, 8]] * 2
    betas = [0.3, 0.5, 0.4, 0.6, 0.45, 0.55]
    return 2, shapes, betas


@pytest.mark.parametrize("kind", ["mixed4d", "uneven3d", "plane"])
@pytest.mark.parametrize("kernel", [O.KERNEL_HEAT_BATH, O.KERNEL_METROPOLIS])
@pytest.mark.parametrize("graph", [1, 0, 2], ids=["graph", "direct", "coop"])
def test_ensemble_is_bit_exact(gpu, kind, kernel, graph):
    dim, shapes, betas = make_case(kind)
    seeds = [100 + r for r in range(len(shapes))]
    sizes = [int(np.prod(s)) * dim for s in shapes]
    start = np.random.default_rng(7).integers(0, 2, sum(sizes)).astype(np.uint8)
    profile = O.PROFILE_F64 if kernel == O.KERNEL_HEAT_BATH else O.PROFILE_F32
    ens = gpu.Ensemble(dim, shapes, betas, seeds, start=start, kernel=kernel, profile=profile, measure_every=2,
                       exchange=True, exchange_seed=11)
    ens.set_launch(128, 2, graph)
    orc = O.Ensemble(dim, shapes, betas, seeds, kernel=kernel, profile=profile, measure_every=2, exchange=True,
                     exchange_seed=11, state=start)
    for chunk in (3, 8, 9):  # crosses block boundaries, uses whole graph blocks
        ens.advance(chunk)
        orc.advance(chunk)
        assert np.array_equal(ens.read(), orc.state)
        assert np.array_equal(ens.betas(), orc.betas_now())
    sums, slots = ens.series()
    assert len(orc.records) == sums.shape[0] == 10
    assert np.array_equal(sums, np.array(orc.records))
    assert np.array_equal(slots, np.array(orc.slots))
    acc, att = ens.exchange_counts()
    assert np.array_equal(acc, orc.accepted[:acc.size])
    assert att.sum() > 0
    live = ens.plaquette_sums()
    off = 0
    for r, s in enumerate(shapes):
        assert live[r] == O.plaquette_sum(s, orc.state[off:off + sizes[r]])
        off += sizes[r]
    ens.close()


def test_launch_shapes_and_tuning_leave_the_chain_unchanged(gpu):
    dim, shapes, betas = make_case("mixed4d")
    seeds = list(range(len(shapes)))
    sizes = [int(np.prod(s)) * dim for s in shapes]
    start = np.random.default_rng(3).integers(0, 2, sum(sizes)).astype(np.uint8)
    outs = []
    for bs, pl, g in ((256, 1, 1), (64, 4, 0), (128, 2, 1), (64, 2, 2)):
        e = gpu.Ensemble(dim, shapes, betas, seeds, start=start, measure_every=5, exchange=True)
        e.set_launch(bs, pl, g)
        e.advance(23)
        outs.append((e.read(), e.betas(), e.series()[0]))
        e.close()
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0])
        assert np.array_equal(o[1], outs[0][1])
        assert np.array_equal(o[2], outs[0][2])
    e = gpu.Ensemble(dim, shapes, betas, seeds, start=start, measure_every=5, exchange=True)
    e.advance(10)
    before = (e.read(), e.betas(), e.series()[0], e.sweeps_done)
    best = e.tune()
    assert best["block_size"] in (64, 128, 256) and best["us_per_sweep"] > 0
    after = (e.read(), e.betas(), e.series()[0], e.sweeps_done)
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    assert np.array_equal(before[2], after[2]) and before[3] == after[3]
    e.advance(13)
    assert np.array_equal(e.read(), outs[0][0])
    e.close()


def test_full_batched_workload_runs_and_agrees_across_launch_shapes(gpu):
    # 48 replicas: L in {8, 12, 16, 20}, 12 each, betas on a uniform grid in
    # [0.40, 0.48]; heat bath; exchange every 10 sweeps
    shapes, betas, seeds = [], [], []
    for L in (8, 12, 16, 20):
        for k in range(12):
            shapes.append([L] * 4)
            betas.append(0.40 + 0.08 * k / 11)
            seeds.append(1000 * L + k)
    sizes = [int(np.prod(s)) * 4 for s in shapes]
    outs = []
    for bs, pl, g in ((256, 1, 1), (128, 4, 0), (256, 1, 2)):
        e = gpu.Ensemble(4, shapes, betas, seeds, measure_every=10, exchange=True, exchange_seed=5)
        e.set_launch(bs, pl, g)
        e.advance(40)
        outs.append((e.read(), e.betas(), e.series()[0], e.exchange_counts()[0]))
        e.close()
    for o in outs[1:]:
        assert np.array_equal(outs[0][0], o[0])
        assert np.array_equal(outs[0][1], o[1])
        assert np.array_equal(outs[0][2], o[2])
        assert np.array_equal(outs[0][3], o[3])
    # mean plaquette per replica from a cold start stays in the ordered phase
    last = outs[0][2][-1]
    n_plaq = np.array([int(np.prod(s)) * 6 for s in shapes])
    mean_p = last / n_plaq
    assert np.all(mean_p > 0.3) and np.all(mean_p <= 1.0)
    # betas stay a permutation of the grid within each size
    b = outs[0][1]
    for g in range(4):
        assert np.allclose(np.sort(b[12 * g:12 * g + 12]), np.sort(betas[12 * g:12 * g + 12]))


@pytest.mark.parametrize("betas", [[0.25, 0.55, 0.85], [0.3, 0.9]])
def test_device_exchange_samples_the_exact_laws(gpu, betas):
    # Heat-bath replicas of a 2x4 Z2 plane (2^16 states, enumerated) with an
    # exchange round after every sweep, on the device: every beta slot
    # samples its exact plaquette law, swaps follow min(1, e^delta) of the
    # recorded sums on the pairs of the round's parity, and each pair's mean
    # acceptance equals the exact <min(1, e^delta)> (tests/exact_z2.py).
    import exact_z2 as X

    R, T = len(betas), 40000
    e = gpu.Ensemble(2, [[2, 4]] * R, betas, [31 + r for r in range(R)], kernel=O.KERNEL_HEAT_BATH,
                     measure_every=1, exchange=True, exchange_seed=17, record_capacity=T)
    e.advance(T)
    sums, slots = e.series()
    acc, att = e.exchange_counts()
    e.close()
    assert sums.shape == (T, R)
    X.check_exchange(sums, slots, betas, X.plaquette_dos(2, 4), O.reduce)
    # pair k is attempted in the rounds t < T with t = k (mod 2)
    assert list(att[:R - 1]) == [(T - k % 2 + 1) // 2 for k in range(R - 1)]
    assert np.all(acc[:R - 1] <= att[:R - 1])


@pytest.mark.parametrize("kind", ["full", "mixed4d", "uneven3d"])
@pytest.mark.parametrize("kernel", [O.KERNEL_HEAT_BATH, O.KERNEL_METROPOLIS])
def test_index_delta_pass_equals_the_coordinate_pass(gpu, monkeypatch, kind, kernel):
    # The index-delta link pass (default where every threshold is at least
    # 1) and the coordinate pass (PLQ_ENS_FAST=0) are independent
    # implementations: states, records, betas and exchange counts agree.
    if kind == "full":
        dim, shapes, betas, seeds = 4, [], [], []
        for L in (8, 12, 16, 20):
            for k in range(12):
                shapes.append([L] * 4)
                betas.append(0.40 + 0.08 * k / 11)
                seeds.append(1000 * L + k)
    else:
        dim, shapes, betas = make_case(kind)
        seeds = [300 + r for r in range(len(shapes))]
    sizes = [int(np.prod(s)) * dim for s in shapes]
    start = np.random.default_rng(17).integers(0, 2, sum(sizes)).astype(np.uint8)
    outs = []
    for fast in ("1", "0"):
        monkeypatch.setenv("PLQ_ENS_FAST", fast)
        e = gpu.Ensemble(dim, shapes, betas, seeds, start=start, kernel=kernel, measure_every=5, exchange=True,
                         exchange_seed=3)
        e.advance(23)
        outs.append((e.read(), e.betas(), e.series()[0], e.exchange_counts()[0]))
        e.close()
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.array_equal(a, b)
