"""The C ABI library on the CPU: it loads, exports every declared symbol, its
host-side threshold tables reproduce the oracle's literal acceptance
arithmetic, and its validation rejects what the reference panics on.  No
kernel runs here."""
import os
import re

import numpy as np
import pytest

import oracle as O
import plaquette_amd as P

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "plaquette_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(plq_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = P.lib()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(P.SIGNATURES)


def test_fold_seed_matches_the_oracle():
    for seed in (0, 1, 12345, 0xDEADBEEFCAFEF00D, (1 << 64) - 1):
        assert P.fold_seed(seed) == O.fold_seed(seed)


CASES = [
    # (dim, kernel, profile, beta, j, h)
    (3, 0, 0, 0.2216546, 1.0, 0.0),
    (3, 1, 0, 0.2216546, 1.0, 0.0),
    (2, 0, 0, 0.4406868, 1.0, 0.0),
    (2, 1, 0, 0.4406868, 1.0, 0.0),
    (2, 0, 1, 0.4406868, 1.0, 0.0),
    (2, 1, 1, 0.4406868, 1.0, 0.0),
    (3, 0, 1, 0.2216546, 1.0, 0.0),
    (3, 1, 1, 0.2216546, 1.0, 0.0),
    (2, 0, 0, 0.7, 0.8, 0.3),
    (4, 1, 0, 0.15, 1.3, -0.2),
    (1, 0, 1, 1.1, 1.0, 0.5),
    (2, 1, 0, 0.0, 1.0, 0.0),
    (3, 0, 0, 3.0, 1.0, 0.0),
]


@pytest.mark.parametrize("case", CASES)
def test_ising_threshold_table_reproduces_the_literal_arithmetic(case):
    dim, kernel, profile, beta, j, h = case
    pd = P.local_desc([4] * dim, P.Model.ISING, 2, kernel, profile, beta, j, h)
    od = O.local_desc([4] * dim, O.MODEL_ISING, 2, kernel, profile, beta, j, h)
    table = P.threshold_table(pd)
    K = 2 * dim + 1
    assert table.size == 2 * K
    rng = np.random.default_rng(dim * 10 + kernel)
    for cur in (0, 1):
        for d in range(K):
            nsum = 2 * dim - 2 * d
            t = int(table[cur * K + d])
            assert 0 <= t <= 1 << 24
            probes = {0, 1, (1 << 24) - 1, (1 << 23)} | {int(x) for x in rng.integers(0, 1 << 24, 40)}
            probes |= {v for v in (t - 1, t, t + 1) if 0 <= v < 1 << 24}
            for m in probes:
                if kernel == 0:
                    want = O.ising_metropolis_accept(od, cur, nsum, m)
                    got = int(m < t)
                else:
                    want = O.ising_heat_bath_choice(od, cur, nsum, m)
                    got = 0 if m < t else 1
                assert got == want, (cur, d, m, t)


def test_bond_probability_profiles_differ_only_in_rounding():
    # the f32 cast of p (gpu_cluster.rs:289,317) moves the threshold by at
    # most one ulp of p at 2^-24 granularity
    import math
    p = 1 - math.exp(-1.0050525)
    assert abs(int(p * 2 ** 24) - int(np.float32(p) * 2 ** 24)) <= 2


def _expect(code, fn, *args):
    with pytest.raises(P.PlaquetteError) as err:
        fn(*args)
    assert err.value.code == code, err.value
    return err.value


def test_odd_extents_are_rejected_before_any_device_work():
    # device.rs:714-719
    e = _expect(P.PLQ_EINVAL, P.ising_gpu_chain, [5, 4], 1.0, 0.0, 0.4, 1, np.zeros(20, np.uint8))
    assert "even" in e.message


def test_bad_descriptors_are_rejected():
    z = np.zeros(16, np.uint8)
    _expect(P.PLQ_EINVAL, P.ising_gpu_chain, [4, 4], 1.0, 0.0, 0.4, 1, np.zeros(15, np.uint8))
    _expect(P.PLQ_EINVAL, P.ising_gpu_chain, [4, 4], 1.0, 0.0, 0.4, 1, np.full(16, 2, np.uint8))
    _expect(P.PLQ_EINVAL, P.potts_gpu_chain, [4, 4], 9, 1.0, None, 0.4, 1, z)
    _expect(P.PLQ_EINVAL, P.gauge_gpu_chain, [4], 1.0, 0.4, 1, np.zeros(4, np.uint8))
    # updater.rs:447-478: negative gap, and the Q = 2 all-clusters forced change
    _expect(P.PLQ_EINVAL, P.GpuClusterChain.new, [4, 4], 3, -1.0, P.Extent.ALL, P.Relabel.REDRAW, 0.5, 1, z)
    _expect(P.PLQ_EINVAL, P.GpuClusterChain.new, [4, 4], 2, 2.0, P.Extent.ALL, P.Relabel.FORCED_CHANGE, 0.5, 1, z)


def test_missing_device_is_an_error_not_a_fallback():
    # Gpu::new() -> None (device.rs:42-77) becomes ENODEV; on a GPU box the
    # same call succeeds, so only check the no-device branch here
    if P.gpu_available():
        pytest.skip("a device is present")
    _expect(P.PLQ_ENODEV, P.ising_gpu_chain, [4, 4], 1.0, 0.0, 0.4, 1, np.zeros(16, np.uint8))
