"""Sample batches and the run-config samplers on the device (GPU).

- ``plq_chain_next_batch`` gives, bit for bit, the samples of as many
  ``plq_chain_next`` calls on a twin chain, for site, link and cluster chains.
- A sampler built from TOML text streams the samples its chain would give by
  hand: the reference's build order (seed, hot start from the seeded stream,
  chain) and phasing (thermalize once, then sweeps_between per sample).
- backend = "cpu" with the checkerboard reproduces the CPU loop with its own
  PCG64 stream continuing after the hot start (the oracle's sequential loop).
- The full stack against the exact 1D correlator, as ising_e2e.rs:332-371
  drives it, and against the 2D area law through the gauge sampler.
"""
import itertools
import math

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _hot(n, q, seed):
    return np.random.default_rng(seed).integers(0, q, n).astype(np.uint8)


@pytest.mark.parametrize("kind", ["ising2d", "ising3d_small", "potts", "gauge4d", "sw", "wolff"])
def test_next_batch_equals_successive_next_calls(gpu, kind):
    def make():
        if kind == "ising2d":
            return gpu.ising_gpu_chain([64, 32], 1.0, 0.1, 0.44, 5, _hot(2048, 2, 1), sweeps_between=3)
        if kind == "ising3d_small":
            return gpu.ising_gpu_chain([8, 8, 8], 1.0, 0.0, 0.22, 5, _hot(512, 2, 2), sweeps_between=2,
                                       kernel=gpu.Kernel.HEAT_BATH)
        if kind == "potts":
            return gpu.potts_gpu_chain([34, 12], 3, 1.0, None, 1.0, 5, _hot(408, 3, 3), sweeps_between=2)
        if kind == "gauge4d":
            return gpu.gauge_gpu_chain([4, 4, 6, 4], 1.0, 0.5, 5, _hot(384 * 4, 2, 4), sweeps_between=1)
        if kind == "sw":
            return gpu.GpuClusterChain.swendsen_wang([33, 17], 3, 1.0, 1.0, 5, _hot(561, 3, 5), sweeps_between=2)
        return gpu.GpuClusterChain.wolff([16, 16], 2, 2.0, 0.44, 5, _hot(256, 2, 6), sweeps_between=3)

    with make() as a, make() as b:
        batch = a.next_batch(7)
        singles = np.stack([next(b) for _ in range(7)])
        assert batch.shape == singles.shape
        assert np.array_equal(batch, singles)
        assert a.sweeps_done == b.sweeps_done
        # and the chains continue in step
        assert np.array_equal(a.next_batch(2), np.stack([next(b), next(b)]))


def _toml(shape, beta, updater, n_samples, thermalize, sweeps_between, seed=20260728, start="hot",
          backend="gpu", schedule=None, extra=""):
    lines = [f"shape = {list(shape)}", "j = 1.0", f"beta = {beta}", f'updater = "{updater}"',
             f'backend = "{backend}"', f"thermalize = {thermalize}", f"sweeps_between = {sweeps_between}",
             f"n_samples = {n_samples}", f"seed = {seed}", f'start = "{start}"']
    if schedule:
        lines.append(f'schedule = "{schedule}"')
    return "\n".join(lines) + "\n" + extra


def test_gpu_sampler_streams_what_the_chain_gives_by_hand(gpu):
    from plaquette_amd.run import GPU_BATCH, IsingRunConfig, IsingSampler

    text = _toml([16, 16], 0.44, "heat_bath", 100, 30, 2, schedule="checkerboard")
    cfg = IsingRunConfig.parse(text)
    with IsingSampler(cfg) as s:
        got = np.stack(list(itertools.islice(s.samples(), GPU_BATCH + 5)))
        start = s.start
    want_start, _ = gpu.hot_start(cfg.seed, 2, 256)
    assert np.array_equal(start, want_start)
    with gpu.ising_gpu_chain([16, 16], 1.0, 0.0, 0.44, cfg.seed, want_start, sweeps_between=2,
                             kernel=gpu.Kernel.HEAT_BATH, profile=gpu.Profile.F32) as c:
        c.advance(30)
        want = np.stack([next(c) for _ in range(GPU_BATCH + 5)])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("kernel_name,start", [("metropolis", "hot"), ("heat_bath", "cold"), ("metropolis", "cold")])
def test_cpu_backend_checkerboard_reproduces_the_cpu_loop(gpu, kernel_name, start):
    from plaquette_amd.run import IsingRunConfig, IsingSampler

    shape, beta, seed = [12, 8], 0.4, 424242
    cfg = IsingRunConfig.parse(_toml(shape, beta, kernel_name, 10, 7, 3, seed=seed, start=start, backend="cpu",
                                     schedule="checkerboard", extra="h = 0.125\n"))
    with IsingSampler(cfg) as s:
        got = np.stack(list(itertools.islice(s.samples(), 10)))
    kernel = O.KERNEL_METROPOLIS if kernel_name == "metropolis" else O.KERNEL_HEAT_BATH
    gen = O.pcg64(seed)
    n = 96
    if start == "hot":
        st, (state, inc) = gpu.hot_start(seed, 2, n)
        gen.state_lo, gen.state_hi = state & (2 ** 64 - 1), state >> 64
        gen.inc_lo, gen.inc_hi = inc & (2 ** 64 - 1), inc >> 64
    else:
        st = np.zeros(n, np.uint8)
    st = O.pmode_ising_sweeps_from(shape, kernel, beta, 1.0, 0.125, gen, st, 7)
    for k in range(10):
        st = O.pmode_ising_sweeps_from(shape, kernel, beta, 1.0, 0.125, gen, st, 3)
        assert np.array_equal(got[k], st), k


def test_samplers_refuse_what_has_no_device_counterpart(gpu):
    from plaquette_amd.run import ConfigError, GaugeRunConfig, GaugeSampler, IsingRunConfig, IsingSampler

    with pytest.raises(ConfigError):
        IsingSampler(IsingRunConfig.parse(_toml([8, 8], 0.4, "metropolis", 1, 0, 1, backend="cpu")))  # random
    with pytest.raises(ConfigError):
        IsingSampler(IsingRunConfig.parse(_toml([8, 8], 0.4, "wolff", 1, 0, 1, backend="cpu")))
    with pytest.raises(ConfigError):
        GaugeSampler(GaugeRunConfig.parse(_toml([4, 4], 0.4, "metropolis", 1, 0, 1, backend="cpu",
                                                schedule="checkerboard")))


@pytest.mark.parametrize("updater", ["heat_bath", "swendsen_wang", "wolff"])
def test_one_dimensional_correlator_through_the_sampler(gpu, updater):
    # ising_e2e.rs:332-371, driven through a run config: <s_0 s_r> against
    # (t^r + t^(N-r)) / (1 + t^N).  The local rule is the checkerboard heat
    # bath (the checkerboard Metropolis is not ergodic in 1D at h = 0).
    from plaquette_amd.run import IsingRunConfig, IsingSampler

    N, beta, samples = 16, 0.4, 4000
    pace = 8 if updater == "wolff" else 2
    text = _toml([N], beta, updater, samples, 500, pace,
                 schedule="checkerboard" if updater == "heat_bath" else None)
    with IsingSampler(IsingRunConfig.parse(text)) as s:
        configs = np.stack(list(itertools.islice(s.samples(), samples)))
    spins = 1 - 2 * configs.astype(np.int64)
    t = math.tanh(beta)
    for r in range(N // 2 + 1):
        measured = float(np.mean(spins * np.roll(spins, -r, axis=1)))
        exact = (t ** r + t ** (N - r)) / (1 + t ** N)
        assert abs(measured - exact) < 0.02, (updater, r, measured, exact)


def test_gauge_sampler_follows_the_area_law(gpu):
    from plaquette_amd.run import GaugeRunConfig, GaugeSampler

    beta = 0.5
    text = _toml([8, 8], beta, "heat_bath", 800, 100, 5, seed=20260805, schedule="checkerboard")
    with GaugeSampler(GaugeRunConfig.parse(text)) as s:
        acc = np.zeros((3, 3))
        for _ in range(800):
            s.chain.advance(5)
            acc += s.chain.wilson_rectangles(2)
    t = math.tanh(beta)
    for r in (1, 2):
        for u in (1, 2):
            assert abs(acc[r][u] / 800 - t ** (r * u)) < 0.02


def test_observables_stream_matches_measuring_the_samples(gpu):
    from plaquette_amd.run import PottsRunConfig, PottsSampler

    text = _toml([32, 32], 1.0, "swendsen_wang", 10, 5, 1)
    cfg = PottsRunConfig.parse(text)
    with PottsSampler(cfg, q=3) as a, PottsSampler(cfg, q=3) as b:
        obs = list(itertools.islice(a.observables(), 6))
        configs = list(itertools.islice(b.samples(), 6))
    for o, c in zip(obs, configs):
        bonds, counts = O.potts_sums([32, 32], 3, c)
        assert o.bond_sum == bonds
        assert [o.counts[k] for k in range(3)] == counts.tolist()
