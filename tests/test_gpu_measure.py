"""The device measurements against the oracle and against exact results (GPU).

Correlators, Polyakov loops and Wilson rectangle tables are computed by the
HIP kernels through the C ABI and compared with the oracle's site-by-site
restatement of the reference on the very configuration the chain holds (read
back through the ABI): integer sums equal and f64 values identical.  At sizes
the oracle cannot walk in seconds, closed forms and invariances take over, and
chains are held to the exact 1D correlator and the exact 2D area law.
"""
import math

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

M, HB = 0, 1


def _hot(n, q, seed):
    return np.random.default_rng(seed).integers(0, q, n).astype(np.uint8)


def _same_floats(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.int64), np.asarray(b, np.float64).view(np.int64))


# ----------------------------------------------------------- correlator ---

SITE_CASES = [
    # (shape, model, q, kernel, sweeps)
    ([8, 6], "ising", 2, M, 0),
    ([8, 6], "ising", 2, M, 3),
    ([64], "ising", 2, HB, 5),
    ([16, 16, 16], "ising", 2, M, 2),
    ([4, 6, 4, 8], "ising", 2, HB, 2),
    ([34, 10], "potts", 3, M, 2),
    ([6, 8, 10], "potts", 5, HB, 2),
    ([32, 32], "potts", 8, M, 1),
    ([256, 256], "ising", 2, M, 4),
    ([96, 2, 4], "potts", 4, HB, 1),
]


@pytest.mark.parametrize("case", SITE_CASES, ids=lambda c: f"{c[1]}{c[2]}-{'x'.join(map(str, c[0]))}-s{c[4]}")
def test_local_chain_correlator_matches_the_oracle(gpu, case):
    shape, model, q, kernel, sweeps = case
    n = int(np.prod(shape))
    st = _hot(n, q, n + q)
    if model == "ising":
        c = gpu.ising_gpu_chain(shape, 1.0, 0.1, 0.4, 11, st, kernel=kernel)
    else:
        c = gpu.potts_gpu_chain(shape, q, 1.0, None, 0.9, 11, st, kernel=kernel)
    with c:
        c.advance(sweeps)
        cfg = c.read()
        sums = c.correlator_sums()
        vals = c.correlator()
    kind = 0 if model == "ising" else 1
    want = O.axis_pair_sums(shape, cfg, kind)
    want_vals = O.ising_correlator(shape, cfg) if model == "ising" else O.potts_correlator(shape, q, cfg)
    assert len(sums) == len(shape)
    for mu in range(len(shape)):
        assert sums[mu].tolist() == want[mu].tolist(), f"axis {mu}"
        assert _same_floats(vals[mu], want_vals[mu]), f"axis {mu}"


CLUSTER_CASES = [
    # (shape, q, convention) -- odd extents and q up to 16
    ([7, 5], 2, "ising"),
    ([33, 9], 2, "potts"),
    ([5, 3, 7], 3, "potts"),
    ([1, 6], 9, "potts"),
    ([3, 3, 3, 5], 16, "potts"),
    ([97], 4, "potts"),
    ([65, 63], 2, "ising"),
]


@pytest.mark.parametrize("case", CLUSTER_CASES, ids=lambda c: f"q{c[1]}-{'x'.join(map(str, c[0]))}-{c[2]}")
def test_cluster_chain_correlator_matches_the_oracle(gpu, case):
    shape, q, conv = case
    n = int(np.prod(shape))
    st = _hot(n, q, 3 * n + q)
    with gpu.GpuClusterChain.swendsen_wang(shape, q, 1.0, 0.6, 5, st) as c:
        c.advance(2)
        cfg = c.read()
        model = gpu.Model.ISING if conv == "ising" else gpu.Model.POTTS
        sums = c.correlator_sums(model)
        vals = c.correlator(model)
    kind = 0 if conv == "ising" else 1
    want = O.axis_pair_sums(shape, cfg, kind)
    want_vals = O.ising_correlator(shape, cfg) if conv == "ising" else O.potts_correlator(shape, q, cfg)
    for mu in range(len(shape)):
        assert sums[mu].tolist() == want[mu].tolist(), f"axis {mu}"
        assert _same_floats(vals[mu], want_vals[mu]), f"axis {mu}"


def test_correlator_at_128_cubed_matches_the_oracle(gpu):
    shape = [128, 128, 128]
    n = 128 ** 3
    with gpu.ising_gpu_chain(shape, 1.0, 0.0, 0.22, 3, _hot(n, 2, 1)) as c:
        c.advance(3)
        cfg = c.read()
        sums = c.correlator_sums()
    want = O.axis_pair_sums(shape, cfg, 0)
    for mu in range(3):
        assert sums[mu].tolist() == want[mu].tolist()


def test_correlator_closed_forms_at_256_cubed(gpu):
    # Aligned: every entry is N.  Checkerboard: (-1)^r N.  Independent of the
    # oracle, and at the size of the bench lattice.
    shape = [256, 256, 256]
    n = 256 ** 3
    x = np.indices((256, 256, 256)).sum(axis=0).reshape(-1) & 1
    with gpu.ising_gpu_chain(shape, 1.0, 0.0, 0.22, 3, np.zeros(n, np.uint8)) as c:
        for row in c.correlator_sums():
            assert row.tolist() == [n] * 129
        c.write(x.astype(np.uint8))
        for row in c.correlator():
            assert row.tolist() == [(-1.0) ** r for r in range(129)]


def test_correlator_rejects_the_wrong_convention(gpu):
    with gpu.ising_gpu_chain([4, 4], 1.0, 0.0, 0.4, 1, np.zeros(16, np.uint8)) as c:
        with pytest.raises(gpu.PlaquetteError):
            c.correlator(gpu.Model.POTTS)
    with gpu.GpuClusterChain.swendsen_wang([4, 4], 3, 1.0, 0.4, 1, np.zeros(16, np.uint8)) as c:
        with pytest.raises(gpu.PlaquetteError):
            c.correlator(gpu.Model.ISING)


# ------------------------------------------------------------- gauge ---

GAUGE_CASES = [
    # (shape, max_side, kernel, sweeps)
    ([4, 4], 2, M, 0),
    ([8, 6], 10, HB, 3),
    ([6, 4, 8], 3, M, 2),
    ([4, 4, 4, 4], 2, HB, 2),
    ([8, 8, 8, 8], 4, M, 1),
    ([66, 4], 2, M, 2),
    ([2, 6, 4, 2], 3, HB, 1),
]


@pytest.mark.parametrize("case", GAUGE_CASES, ids=lambda c: f"{'x'.join(map(str, c[0]))}-m{c[1]}-s{c[3]}")
def test_gauge_measurements_match_the_oracle(gpu, case):
    shape, max_side, kernel, sweeps = case
    D = len(shape)
    n = int(np.prod(shape))
    with gpu.gauge_gpu_chain(shape, 1.0, 0.6, 9, _hot(n * D, 2, n + D), kernel=kernel) as c:
        c.advance(sweeps)
        cfg = c.read()
        table = c.wilson_rectangles(max_side)
        loops = [c.polyakov_loop(d) for d in range(D)]
    assert _same_floats(table, O.wilson_rectangles(shape, cfg, max_side))
    for d in range(D):
        assert _same_floats(loops[d], O.polyakov_loop(shape, cfg, d)), f"direction {d}"


def _gauge_transform(shape, links, sites):
    D = len(shape)
    out = links.copy()
    stride = np.cumprod([1] + list(shape[:-1]))
    for s in sites:
        for mu in range(D):
            out[s * D + mu] ^= 1
            x = (s // stride[mu]) % shape[mu]
            out[(s - x * stride[mu] + ((x - 1) % shape[mu]) * stride[mu]) * D + mu] ^= 1
    return out


def test_gauge_invariance_and_unit_loop_at_16_to_the_fourth(gpu):
    shape = [16, 16, 16, 16]
    D, n = 4, 16 ** 4
    with gpu.gauge_gpu_chain(shape, 1.0, 0.5, 4, _hot(n * D, 2, 2)) as c:
        c.advance(4)
        cfg = c.read()
        table = c.wilson_rectangles(8)
        loops = [c.polyakov_loop(d) for d in range(D)]
        plaq = c.measure().bond_sum
        c.write(_gauge_transform(shape, cfg, range(0, n, 7)))
        assert not np.array_equal(c.read(), cfg)
        assert _same_floats(c.wilson_rectangles(8), table)
        assert [c.polyakov_loop(d) for d in range(D)] == loops
    assert table.shape == (9, 9)
    assert np.array_equal(table, table.T)
    assert table[1][1] == plaq / (n * D * (D - 1) // 2)
    # a cold start: every loop is +1
    with gpu.gauge_gpu_chain(shape, 1.0, 0.5, 4, np.zeros(n * D, np.uint8)) as c:
        assert np.all(c.wilson_rectangles(8) == 1.0)
        assert [c.polyakov_loop(d) for d in range(D)] == [1.0] * D


def test_centre_transformation_flips_one_polyakov_direction(gpu):
    shape = [8, 8, 8]
    D, n = 3, 512
    with gpu.gauge_gpu_chain(shape, 1.0, 0.5, 4, _hot(n * D, 2, 13)) as c:
        cfg = c.read()
        before = [c.polyakov_loop(d) for d in range(D)]
        e0 = c.measure().energy
        flipped = cfg.copy()
        for s in range(n):
            if s % 8 == 2:
                flipped[s * D] ^= 1
        c.write(flipped)
        assert c.measure().energy == e0
        assert c.polyakov_loop(0) == -before[0]
        assert [c.polyakov_loop(1), c.polyakov_loop(2)] == before[1:]


def test_site_chain_measurements_are_refused_on_links_and_back(gpu):
    with gpu.gauge_gpu_chain([4, 4], 1.0, 0.5, 1, np.zeros(32, np.uint8)) as c:
        with pytest.raises(gpu.PlaquetteError):
            c.correlator(gpu.Model.ISING)
        with pytest.raises(gpu.PlaquetteError):
            c.polyakov_loop(2)
    with gpu.ising_gpu_chain([4, 4], 1.0, 0.0, 0.4, 1, np.zeros(16, np.uint8)) as c:
        with pytest.raises(gpu.PlaquetteError):
            c.wilson_rectangles(2)


# ------------------------------------------------------- exact results ---

def _exact_1d(N, beta, r):
    t = math.tanh(beta)
    return (t ** r + t ** (N - r)) / (1 + t ** N)


@pytest.mark.parametrize("updater", ["heat_bath", "swendsen_wang", "wolff"])
def test_one_dimensional_correlator_matches_the_exact_solution(gpu, updater):
    # ising_e2e.rs:332-371: <s_0 s_r> = (t^r + t^(N-r)) / (1 + t^N), t = tanh(beta J).
    # The reference runs its local rule there under the random schedule; the
    # device's checkerboard Metropolis is not ergodic in 1D at h = 0 (the
    # oracle test of the domain-wall imbalance shows why), so the local rule
    # held to the closed form here is the checkerboard heat bath.
    N, beta, samples = 16, 0.4, 4000
    start = np.zeros(N, np.uint8)
    if updater == "heat_bath":
        c = gpu.ising_gpu_chain([N], 1.0, 0.0, beta, 20260805, start, sweeps_between=2, kernel=HB)
    elif updater == "swendsen_wang":
        c = gpu.GpuClusterChain.swendsen_wang([N], 2, 2.0, beta, 20260805, start, sweeps_between=2)
    else:
        c = gpu.GpuClusterChain.wolff([N], 2, 2.0, beta, 20260805, start, sweeps_between=8)
    sums = np.zeros(N // 2 + 1)
    with c:
        c.advance(500)
        step = 2 if updater != "wolff" else 8
        for _ in range(samples):
            c.advance(step)
            sums += c.correlator(gpu.Model.ISING)[0]
    for r in range(N // 2 + 1):
        measured = sums[r] / samples
        assert abs(measured - _exact_1d(N, beta, r)) < 0.02, (updater, r, measured)


def test_two_dimensional_wilson_loops_follow_the_area_law(gpu):
    # gauge/mod.rs:795-830: in 2D every rectangle averages to tanh(beta)^area.
    # The reference samples it under the random schedule; under the device's
    # link checkerboard the Metropolis rule conserves a vortex imbalance in 2D
    # (pinned by an oracle test) and is not ergodic, so the heat bath is held
    # to the area law here.
    beta, samples = 0.5, 800
    with gpu.gauge_gpu_chain([8, 8], 1.0, beta, 20240728, _hot(128, 2, 5), kernel=HB) as c:
        c.advance(100)
        acc = np.zeros((3, 3))
        for _ in range(samples):
            c.advance(5)
            acc += c.wilson_rectangles(2)
    t = math.tanh(beta)
    for r in (1, 2):
        for s in (1, 2):
            assert abs(acc[r][s] / samples - t ** (r * s)) < 0.02, (r, s)
