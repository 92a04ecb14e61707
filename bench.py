"""Benchmark: 3D Ising 256^3 checkerboard Metropolis on one MI355X.

Workload (one "step" = one full checkerboard sweep, both colour passes):
  3D Ising on a 256^3 periodic cubic lattice (16,777,216 int8 spins),
  beta = 0.2216546, J = 1, h = 0, cold start (all +1), energy and
  magnetization reduced on the device every 10 sweeps.  Metropolis is the
  headline; the heat bath is timed separately on the same workload.

Reported:
  value        whole-job site-updates/ns over the timed sweeps (wall clock,
               lattice resident in HBM, device synchronised on both sides);
  roofline     the colour-pass kernel: algorithmic bytes per launch (2 B per
               site-update: int8 read + write; N/2 site-updates per launch)
               over its average launch duration, timed with HIP events on the
               chain's stream, against the 8 TB/s HBM peak;
  cpu_baseline the reference CPU path as written (C restatement of
               LocalUpdate{Metropolis, Checkerboard}: usize neighbour table, u8
               parity table, full scan per colour, f64 exp, PCG64 draws), one
               core, on a bounded sample of the same lattice;
  cpu_baseline_all_cores
               the CPU oracle's keyed-stream checkerboard sweep (the device's own
               semantics) with every colour pass split over all host threads
               (OMP_NUM_THREADS), on the same lattice.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "site-updates/ns (% HBM roofline); tau_int-corrected samples/s vs plain"
HBM_PEAK_GBS = 8000.0
SHAPE = [256, 256, 256]
BETA = 0.2216546
SEED = 20260728


class HipEvents:
    """hipEvent timing on a given stream through libamdhip64."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        rc = self.hip.hipEventCreate(C.byref(e))
        if rc != 0:
            raise RuntimeError(f"hipEventCreate failed ({rc})")
        return e

    def record(self, e, stream):
        rc = self.hip.hipEventRecord(e, C.c_void_p(stream))
        if rc != 0:
            raise RuntimeError(f"hipEventRecord failed ({rc})")

    def elapsed_ms(self, a, b):
        self.hip.hipEventSynchronize(b)
        ms = C.c_float()
        self.hip.hipEventElapsedTime(C.byref(ms), a, b)
        return ms.value


def fused_sweep() -> bool:
    """Whether the 256^3 chain runs one launch per sweep (the fused two-colour
    sweep, the library's default for this lattice; PLQ_FUSED_SWEEP=0 keeps
    two colour-pass launches per sweep)."""
    return os.environ.get("PLQ_FUSED_SWEEP", "1") != "0"


def run_kernel(P, ev, kernel, steps, warmup, kernel_sweeps):
    import numpy as np

    n = int(np.prod(SHAPE))
    chain = P.ising_gpu_chain(SHAPE, 1.0, 0.0, BETA, SEED, np.zeros(n, np.uint8), kernel=kernel)
    # whole job: K sweeps with E, M reduced every 10 sweeps.  The records and
    # events are set up before the warm-up, so the timed sweeps follow the
    # warm-up's synchronisation with no host work between (the recording
    # phase counts from sweep 0: the timed sweeps record as they would after
    # a set-up between warm-up and timing).  The wall-clock region holds the
    # chain's own calls only; the GPU clock of the job comes from K more
    # sweeps between two events (an event record costs the host about 3 us).
    chain.record(10, (2 * steps + warmup) // 10 + 1)
    stream = chain.stream()
    e0, e1 = ev.event(), ev.event()
    chain.advance(warmup)
    chain.sync()
    t0 = time.perf_counter()
    chain.advance(steps)
    chain.sync()
    wall = time.perf_counter() - t0
    ev.record(e0, stream)
    chain.advance(steps)
    ev.record(e1, stream)
    chain.sync()
    gpu_ms = ev.elapsed_ms(e0, e1)
    series = [o for o in chain.series() if warmup < o.sweep <= warmup + steps]  # the timed sweeps' records
    # kernel-only region: the sweep launches alone (no records)
    chain.record(0, 0)
    chain.sync()
    k0, k1 = ev.event(), ev.event()
    ev.record(k0, stream)
    chain.advance(kernel_sweeps)
    ev.record(k1, stream)
    chain.sync()
    launch_ms = ev.elapsed_ms(k0, k1) / ((1 if fused_sweep() else 2) * kernel_sweeps)
    chain.close()
    return {
        "wall_s": wall,
        "gpu_ms": gpu_ms,
        "site_updates_per_ns": steps * n / (wall * 1e9),
        "site_updates_per_ns_gpu_clock": steps * n / (gpu_ms * 1e6),
        "launch_ms": launch_ms,
        "last_energy_per_site": series[-1].energy / n if series else None,
        "last_abs_m_per_site": abs(series[-1].spin_sum) / n if series else None,
        "records": len(series),
    }


def cpu_baseline(seconds_target: float):
    """The reference CPU checkerboard path on one core, bounded sample."""
    import numpy as np

    import oracle as O

    n = int(np.prod(SHAPE))
    st = np.zeros(n, np.uint8)
    st, secs = O.pmode_ising_sweeps(SHAPE, O.KERNEL_METROPOLIS, BETA, 1.0, 0.0, SEED, st, 1)
    sweeps = max(1, int(seconds_target / max(secs, 1e-3)))
    st, secs2 = O.pmode_ising_sweeps(SHAPE, O.KERNEL_METROPOLIS, BETA, 1.0, 0.0, SEED, st, sweeps)
    total_sweeps = 1 + sweeps
    total_secs = secs + secs2
    return {
        "value": total_sweeps * n / (total_secs * 1e9),
        "unit": "site-updates/ns",
        "cores": 1,
        "kind": "port",
        "sample": f"256^3 Ising cold start, beta={BETA}, {total_sweeps} checkerboard Metropolis sweeps "
                  f"({total_secs:.1f} s on one host core; C restatement of updater.rs:77-96,316-324 "
                  f"with the PCG64 stream, table build excluded)",
    }


def cpu_baseline_all_cores(seconds_target: float):
    """The oracle's keyed-stream sweep on every host thread, bounded sample."""
    import numpy as np

    import oracle as O

    n = int(np.prod(SHAPE))
    d = O.local_desc(SHAPE, model=O.MODEL_ISING, q=2, kernel=O.KERNEL_METROPOLIS, beta=BETA, j=1.0, h=0.0,
                     seed=SEED)
    cores = O.max_threads()
    st = np.zeros(n, np.uint8)
    st, secs = O.local_sweeps_mt(d, st, 0, 1, cores)
    sweeps = max(1, int(seconds_target / max(secs, 1e-3)))
    st, secs2 = O.local_sweeps_mt(d, st, 1, sweeps, cores)
    return {
        "value": sweeps * n / (secs2 * 1e9),
        "unit": "site-updates/ns",
        "cores": cores,
        "kind": "port",
        "sample": f"256^3 Ising cold start, beta={BETA}, {sweeps} checkerboard Metropolis sweeps of the keyed "
                  f"stream after one warm-up sweep ({secs2:.1f} s on {cores} host threads, OpenMP over each "
                  f"colour pass; neighbour and parity tables as updater.rs:316-324 uses them, table build excluded)",
    }


def cpu_baseline_ising2d_64(seconds_target: float, start):
    """The reference CPU checkerboard path on 64^2 (PCG64 stream, f64
    acceptance, neighbour and parity tables) on one core, bounded sample,
    from the benchmarked hot start."""
    import oracle as O

    shape = [64, 64]
    n = 4096
    st = start.copy()
    st, secs = O.pmode_ising_sweeps(shape, O.KERNEL_METROPOLIS, 0.4406868, 1.0, 0.0, SEED, st, 100)
    sweeps = max(100, int(100 * seconds_target / max(secs, 1e-6)))
    st, secs2 = O.pmode_ising_sweeps(shape, O.KERNEL_METROPOLIS, 0.4406868, 1.0, 0.0, SEED, st, sweeps)
    return {
        "value": sweeps * n / (secs2 * 1e9), "unit": "site-updates/ns", "cores": 1, "kind": "port",
        "sample": f"64^2 Ising at beta_c, {sweeps} checkerboard Metropolis sweeps ({secs2:.1f} s on one host core; "
                  f"C restatement of updater.rs:77-96,316-324 with the PCG64 stream)",
    }


def cpu_baseline_potts_sw(seconds_target: float, start):
    """The oracle's sequential Swendsen-Wang sweep at 4096^2 on one core: the
    reference CPU ClusterUpdate (updater.rs:541-613) -- bond draw, union-find
    forest with minimum-site roots (cluster.rs:44-227), relabel per cluster --
    driven by the keyed stream; bounded sample of whole sweeps from the
    benchmarked hot start."""
    import oracle as O

    shape = [4096, 4096]
    n = 4096 * 4096
    beta = math.log(1 + math.sqrt(3))
    d = O.cluster_desc(shape, q=3, beta=beta, gap=1.0, seed=SEED)
    st = start.copy()
    t0 = time.perf_counter()
    st, _ = O.cluster_sweeps(d, st, 0, 1)
    one = time.perf_counter() - t0
    sweeps = max(1, int(seconds_target / max(one, 1e-3)))
    t0 = time.perf_counter()
    st, _ = O.cluster_sweeps(d, st, 1, sweeps)
    secs = time.perf_counter() - t0
    return {
        "value": sweeps * n / (secs * 1e9), "unit": "site-updates/ns", "cores": 1, "kind": "port",
        "sample": f"4096^2 q=3 Potts at beta_c, {sweeps} Swendsen-Wang sweeps after one ({secs:.1f} s on one host "
                  f"core; C restatement of updater.rs:541-613 and cluster.rs:44-227)",
    }


def cpu_baseline_z2_batch(seconds_target: float):
    """The oracle's sequential Z2 link heat bath (staple table, f64 weights;
    updater.rs:109-167, 325-343) over the 48-replica workload on one core.
    Per replica size the table build is taken out by timing one sweep and
    k + 1 sweeps; the rate is links per second over all 48 replicas."""
    import numpy as np

    import oracle as O

    shapes, betas, seeds = z2_batch_shapes()
    per_size = {}
    budget = seconds_target / 4
    for L in (8, 12, 16, 20):
        shape = [L] * 4
        nl = L ** 4 * 4
        d = O.local_desc(shape, model=O.MODEL_Z2, kernel=O.KERNEL_HEAT_BATH, beta=0.44, seed=SEED)
        st = np.zeros(nl, np.uint8)
        t0 = time.perf_counter()
        st = O.local_sweeps(d, st, 0, 1)
        one = time.perf_counter() - t0
        k = max(2, int(budget / max(one, 1e-4)))
        t0 = time.perf_counter()
        st = O.local_sweeps(d, st, 0, k + 1)
        many = time.perf_counter() - t0
        per_size[L] = max(many - one, 1e-9) / k  # seconds per sweep of one replica, table build excluded
    links = sum(int(np.prod(s)) * 4 for s in shapes)
    secs_per_sweep = sum(per_size[s[0]] for s in shapes)
    return {
        "value": links / (secs_per_sweep * 1e9), "unit": "link-updates/ns", "cores": 1, "kind": "port",
        "sample": "48-replica 4D Z2 workload, link heat bath at beta 0.44: one replica per size timed over "
                  + ", ".join(f"L={L}: {per_size[L] * 1e3:.1f} ms/sweep" for L in per_size)
                  + " on one host core (C restatement of updater.rs:109-167,325-343 with the staple table, "
                  "table build excluded), summed over the 12 replicas of each size",
    }


def read_profile_value(name: str, key: str):
    """One `key: value` line of a committed rocprofv3 summary under profiles/."""
    path = os.path.join(REPO, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            if k.strip() == key:
                try:
                    return float(v.strip())
                except ValueError:
                    return None
    return None


def kernel_files():
    """The committed rocprofv3 summaries of the launch that dominates the
    3D job: the fused sweep, or the colour pass with PLQ_FUSED_SWEEP=0."""
    if fused_sweep():
        return "ising3d_256_sweep_traffic.txt", "ising3d_256_sweep_counters.txt"
    return "ising3d_256_pass_traffic.txt", "ising3d_256_pass_counters.txt"


def read_traffic():
    """HBM bytes per launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    return read_profile_value(kernel_files()[0], "hbm_bytes_per_launch")


def valu_roof(launch_ms: float):
    name = kernel_files()[1]
    lane_ops = read_profile_value(name, "valu_lane_ops_per_launch")
    peak = read_profile_value(name, "valu_peak_lane_ops_per_s")
    per_site = read_profile_value(name, "valu_insts_per_site_update")
    if not lane_ops or not peak:
        return None
    achieved = lane_ops / (launch_ms * 1e-3)
    return {"achieved": round(achieved / 1e12, 2), "peak": round(peak / 1e12, 2), "unit": "T lane-ops/s",
            "frac": round(achieved / peak, 4), "valu_insts_per_site_update": per_site,
            "source": "profiles/" + name}


def profile_passes(sweeps: int, copy_mib: int) -> None:
    """A short fixed workload for hardware-counter runs: `sweeps` Metropolis
    sweeps of the 256^3 lattice (one fused launch each, or two colour-pass
    launches with PLQ_FUSED_SWEEP=0; no records), then one device-to-device
    copy of `copy_mib` MiB whose FETCH_SIZE / WRITE_SIZE calibrate the
    counters against known bytes."""
    import numpy as np

    import plaquette_amd as P

    n = int(np.prod(SHAPE))
    chain = P.ising_gpu_chain(SHAPE, 1.0, 0.0, BETA, SEED, np.zeros(n, np.uint8))
    chain.advance(2)
    chain.sync()
    chain.advance(sweeps)
    chain.sync()
    chain.close()
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    hip.hipFree.argtypes = [C.c_void_p]
    nbytes = copy_mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), nbytes) or hip.hipMalloc(C.byref(b), nbytes):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemcpy(b, a, nbytes, 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    per = 1 if fused_sweep() else 2
    print(json.dumps({"profile": "sweeps", "sweeps": sweeps, "launches": per * sweeps,
                      "copy_bytes": nbytes, "site_updates_per_launch": n // per}))


def timed(P, ev, stream_of, run):
    """Wall and event time of run() bracketed by device synchronisation."""
    e0, e1 = ev.event(), ev.event()
    t0 = time.perf_counter()
    ev.record(e0, stream_of())
    run()
    ev.record(e1, stream_of())
    ms = ev.elapsed_ms(e0, e1)
    return time.perf_counter() - t0, ms


def bench_potts_sw(P, ev, steps, warmup, cpu_seconds=None):
    """2D q=3 Potts 4096^2 Swendsen-Wang at beta_c = ln(1 + sqrt 3); order
    parameter (label counts) recorded every sweep; cluster-size histogram of
    the last sweep, and separately the rate with the histogram of every
    sweep kept on the device.  10 B algorithmic per site and sweep (labels read and
    written, 4-byte cluster id written and read)."""
    import math

    import numpy as np

    shape = [4096, 4096]
    n = 4096 * 4096
    beta = math.log(1 + math.sqrt(3))
    # the run config's hot start: the seeded stream's draws (configuration.rs:80-87)
    start = P.hot_start(SEED, 3, n)[0]
    c = P.GpuClusterChain.swendsen_wang(shape, 3, 1.0, beta, SEED, start)
    c.advance(warmup)
    c.sync()
    c.record(1, steps + 1)
    wall, ms = timed(P, ev, c.stream, lambda: (c.advance(steps), c.sync()))
    series = c.series()
    t0 = time.perf_counter()
    stats = c.stats()
    stats_s = time.perf_counter() - t0
    # the configuration's full per-sweep work: a sweep, the order parameter
    # and the cluster-size histogram of its partition, every sweep, kept on
    # the device (plq_chain_record_clusters)
    hist_steps = 100
    c.record(1, hist_steps)
    c.record_clusters(True)
    c.advance(2)
    c.sync()
    c.record(1, hist_steps)
    hist_wall, _ = timed(P, ev, c.stream, lambda: (c.advance(hist_steps), c.sync()))
    hist_series = c.cluster_series()
    c.close()
    last = series[-1]
    counts = list(last.counts[:3])
    order = (3 * max(counts) / n - 1) / 2
    rate = steps * n / (wall * 1e9)
    return {
        "metric": METRIC, "value": round(rate, 2), "unit": "site-updates/ns", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round(wall * 1e3 / steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "2D q=3 Potts 4096^2 Swendsen-Wang, beta=ln(1+sqrt 3), hot start, order parameter "
                               "every sweep", "lattice": shape, "beta": beta},
        "roofline": {"bound": "hbm", "achieved": round(rate * 10, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(rate * 10 / HBM_PEAK_GBS, 4), "traffic": None,
                     "note": "10 B algorithmic per site-sweep over the whole sweep (all kernels)"},
        "last_energy_per_site": last.energy / n, "last_order_parameter": order,
        "clusters": stats.n_clusters, "largest_cluster": stats.largest, "stats_ms": round(stats_s * 1e3, 2),
        "histogram_every_sweep": {"sweeps": hist_steps, "us_per_sweep": round(hist_wall * 1e6 / hist_steps, 1),
                                  "site_updates_per_ns": round(hist_steps * n / (hist_wall * 1e9), 2),
                                  "histograms_kept": len(hist_series),
                                  "last_clusters": hist_series[-1].n_clusters if hist_series else None},
        "cpu_baseline": cpu_baseline_potts_sw(cpu_seconds, start) if cpu_seconds else None,
    }


def z2_batch_shapes():
    shapes, betas, seeds = [], [], []
    for L in (8, 12, 16, 20):
        for k in range(12):
            shapes.append([L] * 4)
            betas.append(0.40 + 0.08 * k / 11)
            seeds.append(SEED + 1000 * L + k)
    return shapes, betas, seeds


def cache_copy_gbs(ev, nbytes: int, reps: int) -> float:
    """Device-to-device copy rate (bytes read plus written per second, GB/s)
    of an nbytes buffer copied reps times back to back on the null stream:
    the measured ceiling for a working set that stays in the caches."""
    hip = ev.hip
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    hip.hipFree.argtypes = [C.c_void_p]
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), nbytes) or hip.hipMalloc(C.byref(b), nbytes):
        raise RuntimeError("hipMalloc failed")
    try:
        for _ in range(10):
            hip.hipMemcpyAsync(b, a, nbytes, 3, None)  # hipMemcpyDeviceToDevice
        e0, e1 = ev.event(), ev.event()
        ev.record(e0, 0)
        for _ in range(reps):
            hip.hipMemcpyAsync(b, a, nbytes, 3, None)
        ev.record(e1, 0)
        ms = ev.elapsed_ms(e0, e1)
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return 2.0 * nbytes * reps / (ms * 1e-3) / 1e9


def bench_z2_batch(P, ev, steps, warmup, cpu_seconds=None):
    """48 replicas of 4D Z2 (L = 8, 12, 16, 20; 12 each), link heat bath, beta
    grid of 12 points in [0.40, 0.48] per size, replica exchange every 10
    sweeps, cold start; launch shape picked by the autotuner."""
    import numpy as np

    shapes, betas, seeds = z2_batch_shapes()
    links = sum(int(np.prod(s)) * 4 for s in shapes)
    e = P.Ensemble(4, shapes, betas, seeds, kernel=P.Kernel.HEAT_BATH, measure_every=10, exchange=True,
                   exchange_seed=SEED, record_capacity=steps // 10 + 8)
    tuned = e.tune()
    e.advance(warmup)
    e.sync()
    wall, ms = timed(P, ev, e.stream, lambda: (e.advance(steps), e.sync()))
    sums, _ = e.series()
    acc, att = e.exchange_counts()
    # the same sweeps with every colour pass of a 10-sweep block in one
    # cooperative launch (grid barriers between the passes), for comparison
    e.set_launch(tuned["block_size"], tuned["per_lane"], 2)
    e.advance(warmup)
    e.sync()
    coop_wall, _ = timed(P, ev, e.stream, lambda: (e.advance(steps), e.sync()))
    e.close()
    n_plaq = np.array([int(np.prod(s)) * 6 for s in shapes])
    rate = steps * links / (wall * 1e9)
    cache_gbs = cache_copy_gbs(ev, links, 500)
    return {
        "metric": METRIC, "value": round(rate, 2), "unit": "link-updates/ns", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round(wall * 1e3 / steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "batched 4D Z2: 48 replicas, L in {8,12,16,20} x 12, link heat bath, beta grid "
                               "[0.40,0.48], replica exchange every 10 sweeps, cold start", "links": links},
        "roofline": {"bound": "hbm", "achieved": round(rate * 2, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(rate * 2 / HBM_PEAK_GBS, 4), "traffic": None,
                     "note": "2 B algorithmic per link-update; the 12 MB of links stay cache resident"},
        "roofline_cache": {"achieved": round(rate * 2, 1), "peak": round(cache_gbs, 1), "unit": "GB/s",
                           "frac": round(rate * 2 / cache_gbs, 4),
                           "copy_us": round(2.0 * links / (cache_gbs * 1e3), 2),
                           "note": "peak: measured device-to-device copy of a buffer the size of the links, "
                                   "500 times back to back (bytes read plus written); one launch per copy, "
                                   "so launch time is in it, as it is in the eight passes of a sweep"},
        "tuned_launch": tuned,
        "cooperative_launch": {"us_per_sweep": round(coop_wall * 1e6 / steps, 2),
                               "link_updates_per_ns": round(steps * links / (coop_wall * 1e9), 2),
                               "note": "every pass of a 10-sweep block in one cooperative launch, grid barriers "
                                       "between the passes"},
        "mean_plaquette_last": [round(float(x), 5) for x in (sums[-1] / n_plaq)] if len(sums) else None,
        "exchange_acceptance": round(float(acc.sum()) / max(1, float(att.sum())), 4),
        "cpu_baseline": cpu_baseline_z2_batch(cpu_seconds) if cpu_seconds else None,
    }


def bench_ising2d_64(P, ev, steps, warmup, cpu_seconds=None):
    """2D Ising 64^2 at beta_c, checkerboard Metropolis, hot start, E and M
    recorded every sweep (latency bound: 4,096 sites)."""
    import numpy as np

    shape = [64, 64]
    n = 4096
    # the run config's hot start: the seeded stream's draws (configuration.rs:80-87, ising/run_config.rs:237-251)
    start = P.hot_start(SEED, 2, n)[0]
    c = P.ising_gpu_chain(shape, 1.0, 0.0, 0.4406868, SEED, start)
    c.advance(warmup)
    c.sync()
    c.record(1, steps + 1)
    wall, ms = timed(P, ev, c.stream, lambda: (c.advance(steps), c.sync()))
    series = c.series()
    c.close()
    e = np.array([o.energy / n for o in series])
    m = np.array([abs(o.spin_sum) / n for o in series])
    rate = steps * n / (wall * 1e9)
    return {
        "metric": METRIC, "value": round(rate, 4), "unit": "site-updates/ns", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round(wall * 1e3 / steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "2D Ising 64^2, beta=0.4406868, checkerboard Metropolis, hot start, E and |M| "
                               "every sweep", "lattice": shape},
        "mean_energy_per_site": float(e.mean()), "mean_abs_m_per_site": float(m.mean()),
        "cpu_baseline": cpu_baseline_ising2d_64(cpu_seconds, start) if cpu_seconds else None,
    }


def bench_cluster_small(P, ev, steps, warmup):
    """Small-lattice cluster chains: 2D q=3 Potts 64^2 Swendsen-Wang at beta_c
    and 2D Ising 64^2 Wolff at beta_c, hot starts, a record every sweep (one
    workgroup per chain, the whole lattice in LDS)."""
    import numpy as np

    shape, n = [64, 64], 4096
    out = {}
    for name, q, gap, beta, make in (("sw_potts3", 3, 1.0, 1.0050525, P.GpuClusterChain.swendsen_wang),
                                     ("wolff_ising", 2, 2.0, 0.4406868, P.GpuClusterChain.wolff)):
        start = np.random.default_rng(SEED).integers(0, q, n).astype(np.uint8)
        c = make(shape, q, gap, beta, SEED, start)
        c.record(1, steps + 1)
        c.advance(warmup)
        c.sync()
        wall, ms = timed(P, ev, c.stream, lambda: (c.advance(steps), c.sync()))
        c.close()
        out[name] = {"us_per_sweep": round(wall * 1e6 / steps, 3), "site_updates_per_ns": round(steps * n / (wall * 1e9), 4)}
    return {
        "metric": "site-updates/ns, small-lattice cluster chains", "value": out["sw_potts3"]["site_updates_per_ns"],
        "unit": "site-updates/ns", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(out["sw_potts3"]["us_per_sweep"] / 1e3, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "64^2 q=3 Potts Swendsen-Wang (value) and 64^2 Ising Wolff, beta_c, hot start, "
                               "a record every sweep", "lattice": shape},
        "chains": out,
    }


def bench_wolff(P, ev, steps, warmup):
    """Wolff moves on a 2D Ising 4096^2 lattice (gap 2J) at beta 0.3, 0.4,
    beta_c and 0.6, each from a configuration equilibrated by 300
    Swendsen-Wang sweeps: the reference's device move (the full
    decomposition, PLQ_WOLFF_GROW=0) against the default adaptive growth of
    the seed's cluster (PLQ_WOLFF_GROW=1), the same chain either way.  One
    step is one Wolff move."""
    import numpy as np

    L = 4096
    n = L * L
    out = {}
    for beta in (0.3, 0.4, 0.4406868, 0.6):
        start = np.random.default_rng(SEED).integers(0, 2, n).astype(np.uint8)
        sw = P.GpuClusterChain.swendsen_wang([L, L], 2, 2.0, beta, SEED, start)
        sw.advance(300)
        st = sw.read()
        sw.close()
        row = {}
        for name, mode in (("decomposition", "0"), ("adaptive_growth", "1")):
            os.environ["PLQ_WOLFF_GROW"] = mode
            c = P.GpuClusterChain.wolff([L, L], 2, 2.0, beta, SEED + 1, st)
            c.advance(warmup)
            c.sync()
            g0, f0 = c.wolff_counts()
            wall, _ = timed(P, ev, c.stream, lambda: (c.advance(steps), c.sync()))
            g1, f1 = c.wolff_counts()
            c.close()
            row[name] = {"us_per_move": round(wall * 1e6 / steps, 2), "grown": g1 - g0, "decomposed": f1 - f0}
        os.environ.pop("PLQ_WOLFF_GROW", None)
        row["speedup"] = round(row["decomposition"]["us_per_move"] / row["adaptive_growth"]["us_per_move"], 3)
        out[f"beta_{beta}"] = row
    head = out["beta_0.4"]["adaptive_growth"]
    return {
        "metric": "Wolff moves/s, 2D Ising 4096^2 at beta 0.4, adaptive growth", "value": round(1e6 / head["us_per_move"], 1),
        "unit": "moves/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(head["us_per_move"] / 1e3, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic (Swendsen-Wang equilibrated configurations)",
        "config": {"workload": "2D Ising 4096^2 Wolff moves, gap 2J, beta 0.3 / 0.4 / beta_c / 0.6", "lattice": [L, L]},
        "by_beta": out,
    }


def bench_measure(P, ev, reps):
    """The device measurements: the Ising correlator of a 256^3 chain (387
    axis-displacement pair sums) and the Wilson table of a 4D Z2 chain at
    32^4 (sides 1..16, 3,072 rectangle classes), each timed per call with the
    pack, launches and readback."""
    import numpy as np

    def per_call(f):
        f()
        t = time.perf_counter()
        for _ in range(reps):
            f()
        return (time.perf_counter() - t) / reps

    n3 = 256 ** 3
    c = P.ising_gpu_chain([256] * 3, 1.0, 0.0, BETA, SEED, np.random.default_rng(SEED).integers(0, 2, n3).astype(np.uint8))
    corr_s = per_call(c.correlator)
    c.close()
    n4 = 32 ** 4
    g = P.gauge_gpu_chain([32] * 4, 1.0, 0.5, SEED, np.random.default_rng(SEED).integers(0, 2, 4 * n4).astype(np.uint8))
    wil_s = per_call(lambda: g.wilson_rectangles(16))
    pol_s = per_call(lambda: g.polyakov_loop(3))
    g.close()
    pairs = 3 * (256 // 2 + 1) * n3
    return {
        "metric": "site pairs/ns, device correlator", "value": round(pairs / (corr_s * 1e9), 2),
        "unit": "site pairs/ns", "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": round(corr_s * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "Ising correlator of a 256^3 random configuration; Wilson table (side 16) and a "
                               "Polyakov loop of a 32^4 Z2 configuration"},
        "correlator_ms": round(corr_s * 1e3, 4), "wilson_ms": round(wil_s * 1e3, 4), "polyakov_ms": round(pol_s * 1e3, 4),
    }


BLOCK_FLOP = 1253376  # network FLOPs per 8x8 block: 64 x 32 x 9 x 2 (conv1) + 64 x 32 x 288 x 2 (conv2) + 64 x 288 x 2 (conv3)


def bench_cpu_stream(P, ev, steps, warmup):
    """3D Ising 256^3 at beta = 0.2216546, checkerboard Metropolis with the
    CPU path's own PCG64 stream (Profile.CPU_STREAM): the same configurations
    as the reference CPU run with the same seed, on the device."""
    import numpy as np

    n = int(np.prod(SHAPE))
    c = P.ising_gpu_chain(SHAPE, 1.0, 0.0, BETA, SEED, np.zeros(n, np.uint8), profile=P.Profile.CPU_STREAM)
    c.advance(warmup)
    c.sync()
    wall, ms = timed(P, ev, c.stream, lambda: (c.advance(steps), c.sync()))
    c.close()
    rate = steps * n / (wall * 1e9)
    return {
        "metric": "site-updates/ns, reference CPU stream reproduced", "value": round(rate, 3),
        "unit": "site-updates/ns", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(wall * 1e3 / steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "3D Ising 256^3, checkerboard Metropolis, beta=0.2216546, cold start, PCG64 "
                               "seed_from_u64 stream in CPU site order", "lattice": SHAPE},
    }


W3_SLICE = slice(9568, 9857)  # the last layer's weights and bias: the logits are linear in them


def calibrate_sharpness(P, chain, net, sweeps=100):
    """Scale the last layer (logits x alpha) to the alpha of highest block
    acceptance on `chain`, from {1, 1.25, 1.5, 1.75, 2, 2.5}.  The proposal
    stays exact for any alpha: the MH ratio uses the logits proposed with."""
    base = net.params
    best = (0.0, 1.0)
    for alpha in (1.0, 1.25, 1.5, 1.75, 2.0, 2.5):
        p = base.copy()
        p[W3_SLICE] *= alpha
        net.params = p
        a0, p0, _ = chain.block_stats()
        chain.block_sweeps(net, sweeps)
        chain.sync()
        a1, p1, _ = chain.block_stats()
        best = max(best, ((a1 - a0) / (p1 - p0), alpha))
    p = base.copy()
    p[W3_SLICE] *= best[1]
    net.params = p
    return {"alpha": best[1], "acceptance": round(best[0], 5)}


def compare_chains(P, ev, stats, L, beta, st, net, steps, every, seed):
    """Plain (one checkerboard sweep per step) and mixed (one block sweep of
    both colours, then one checkerboard sweep) chains from the same
    configuration, `steps` steps each with E and |M| recorded every `every`
    steps.  tau_int (in steps) from Sokal's window (statistics.rs:83-137);
    samples/s = steps/s / (2 tau_int); a window counts as converged when the
    series spans at least 50 tau_int (Sokal's rule of thumb)."""
    import numpy as np

    n = L * L
    out = {}
    for mode in ("plain", "mixed"):
        ch = P.ising_gpu_chain([L, L], 1.0, 0.0, beta, seed, st)
        ch.record(every, steps // every + 1)
        ch.sync()
        if mode == "plain":
            wall, _ = timed(P, ev, ch.stream, lambda: (ch.advance(steps), ch.sync()))
        else:
            wall, _ = timed(P, ev, ch.stream, lambda: (ch.advance_with_blocks(net, steps, 1), ch.sync()))
        ser = ch.series()
        acc, prop, _ = ch.block_stats()
        ch.close()
        res = {"steps": steps, "steps_per_s": round(steps / wall, 1), "record_every": every,
               "block_acceptance": round(acc / prop, 5) if prop else None}
        for name, series in (("abs_m", [abs(o.spin_sum) / n for o in ser]), ("energy", [o.energy / n for o in ser])):
            est = stats.reduce(np.array(series))
            tau = est.tau_int * every
            res[name] = {"mean": est.mean, "stderr": est.stderr, "tau_int_steps": round(tau, 2),
                         "n_eff": round(est.n_eff, 1), "samples_per_s": round(steps / wall / (2.0 * tau), 4),
                         "window_converged": bool(len(series) >= 50 * est.tau_int)}
        out[mode] = res
    for name in ("abs_m", "energy"):
        out[f"ratio_vs_plain_{name}"] = round(out["mixed"][name]["samples_per_s"] / out["plain"][name]["samples_per_s"], 4)
        d = out["mixed"][name]["mean"] - out["plain"][name]["mean"]
        err = math.hypot(out["mixed"][name]["stderr"], out["plain"][name]["stderr"])
        out[f"agreement_sigma_{name}"] = round(abs(d) / err, 2) if err > 0 else None
    return out


def learned_small(P, ev, stats):
    """The same comparison on 64^2 at beta_c: train 2,000 steps on the
    chain's configurations, calibrate the logit scale, then plain and mixed
    chains of 400,000 steps each (tau_int of |M| is about 180 sweeps)."""
    import numpy as np

    L, beta = 64, 0.4406868
    n = L * L
    start = np.random.default_rng(SEED + 2).integers(0, 2, n).astype(np.uint8)
    c = P.ising_gpu_chain([L, L], 1.0, 0.0, beta, SEED + 2, start)
    c.advance(5000)
    net = P.BlockNet(SEED + 2)
    net.train(c, 2000, 5, 3e-3)
    cal = calibrate_sharpness(P, c, net)
    st = c.read()
    c.close()
    out = compare_chains(P, ev, stats, L, beta, st, net, 400000, 1, SEED + 3)
    out["calibration"] = cal
    net.close()
    return out


def bench_learned(P, ev, steps=1000000, train=20000, with_small=True):
    """2D Ising 512^2 at beta_c: the learned block proposal against the plain
    checkerboard.  A chain from a hot start runs 2,000 sweeps, then the
    network trains `train` (20,000) Adam steps, one per configuration taken
    every 5 sweeps (20,000 configurations of the plain sampler), and the
    logit scale is calibrated on block acceptance.  From the trained chain's
    configuration a plain chain and a mixed chain each run `steps`
    (1,000,000) steps with E and |M| recorded every 10 steps
    (compare_chains).  Then the same comparison at 64^2 (learned_small)
    unless with_small is false.  Smaller counts, or no 64^2 part, only
    shorten a profiled run."""
    import numpy as np

    from plaquette_amd import statistics as stats

    L, beta = 512, 0.4406868
    n = L * L
    start = np.random.default_rng(SEED).integers(0, 2, n).astype(np.uint8)
    c = P.ising_gpu_chain([L, L], 1.0, 0.0, beta, SEED, start)
    c.advance(2000)
    net = P.BlockNet(SEED)
    wall_train, ms_train = timed(P, ev, c.stream, lambda: net.train(c, train, 5, 3e-3))
    losses = net.train(c, 1, 5, 3e-3)
    cal = calibrate_sharpness(P, c, net)
    st = c.read()
    # the block sweep alone, by events on its stream
    k = 50
    c.block_sweeps(net, 5)
    c.sync()
    _, ms_blocks = timed(P, ev, c.stream, lambda: (c.block_sweeps(net, k), c.sync()))
    block_us = ms_blocks * 1e3 / k
    c.close()
    runs = compare_chains(P, ev, stats, L, beta, st, net, steps, 10, SEED + 1)
    net.close()
    small = learned_small(P, ev, stats) if with_small else None
    flop = (n // 64) * BLOCK_FLOP
    tflops = flop / (block_us * 1e-6) / 1e12
    mixed = runs["mixed"]
    return {
        "metric": "tau_int-corrected samples/s of |M|, learned block proposal vs plain checkerboard",
        "value": mixed["abs_m"]["samples_per_s"], "unit": "samples/s", "n_gpus": 1, "steps": steps,
        "warmup": 2000, "ms_per_step": round(1e3 / mixed["steps_per_s"], 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 network, fp32 accumulate; u8 states",
        "data": "synthetic (chain configurations)",
        "config": {"workload": "2D Ising 512^2 at beta_c = 0.4406868, 4,096 blocks of 8x8, 3-layer 3x3 CNN with "
                               f"32 channels (middle layer dilated by 2), trained on {train:,} configurations",
                   "lattice": [L, L]},
        "ratio_vs_plain": runs["ratio_vs_plain_abs_m"], "comparison_512": runs, "comparison_64": small,
        "training": {"steps": train, "wall_s": round(wall_train, 3), "final_loss": float(losses[-1]),
                     "calibration": cal},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(tflops / 2500.0, 5), "traffic": None,
                     "kernel": "block_pass_kernel (both block colours)", "avg_sweep_us": round(block_us, 2)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--profile", action="store_true",
                    help="run the fixed counter-collection workload instead of the benchmark")
    ap.add_argument("--workload", default="ising3d",
                    choices=["ising3d", "potts-sw", "z2-batch", "ising2d-64", "learned", "cpu-stream", "cluster-small",
                             "measure", "wolff"],
                    help="ising3d is the headline line; the others time the remaining configurations")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--kernel-sweeps", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--learned-steps", type=int, default=1000000,
                    help="steps of each compared chain in the learned workload")
    ap.add_argument("--learned-train", type=int, default=20000,
                    help="training configurations in the learned workload")
    ap.add_argument("--learned-no-64", action="store_true",
                    help="leave out the 64^2 comparison of the learned workload")
    args = ap.parse_args()
    if args.profile:
        profile_passes(10, 256)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on these single-GPU boxes "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import plaquette_amd as P

    if not P.gpu_available():
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    ev = HipEvents()
    if args.workload == "potts-sw":
        print(json.dumps(bench_potts_sw(P, ev, 200, 10, None if args.no_cpu else args.cpu_seconds)))
        return
    if args.workload == "z2-batch":
        print(json.dumps(bench_z2_batch(P, ev, 500, 20, None if args.no_cpu else args.cpu_seconds)))
        return
    if args.workload == "cpu-stream":
        print(json.dumps(bench_cpu_stream(P, ev, 50, 5)))
        return
    if args.workload == "learned":
        print(json.dumps(bench_learned(P, ev, args.learned_steps, args.learned_train, not args.learned_no_64)))
        return
    if args.workload == "ising2d-64":
        print(json.dumps(bench_ising2d_64(P, ev, 10000, 1000, None if args.no_cpu else args.cpu_seconds)))
        return
    if args.workload == "wolff":
        print(json.dumps(bench_wolff(P, ev, 300, 20)))
        return
    if args.workload == "cluster-small":
        print(json.dumps(bench_cluster_small(P, ev, 10000, 500)))
        return
    if args.workload == "measure":
        print(json.dumps(bench_measure(P, ev, 20)))
        return
    n = int(np.prod(SHAPE))
    metro = run_kernel(P, ev, P.Kernel.METROPOLIS, args.steps, args.warmup, args.kernel_sweeps)
    hb = run_kernel(P, ev, P.Kernel.HEAT_BATH, args.steps, args.warmup, args.kernel_sweeps)

    # int8 read + write per site-update: N updates per fused sweep launch, N/2 per colour pass
    updates_per_launch = n if fused_sweep() else n // 2
    bytes_per_launch = 2.0 * updates_per_launch
    achieved = bytes_per_launch / (metro["launch_ms"] * 1e-3) / 1e9
    traffic = read_traffic()
    out = {
        "metric": METRIC,
        "value": round(metro["site_updates_per_ns"], 2),
        "unit": "site-updates/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(metro["wall_s"] * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": "3D Ising 256^3 periodic, checkerboard Metropolis, beta=0.2216546, cold start, "
                        "E and M reduced on device every 10 sweeps",
            "lattice": SHAPE,
            "sites": n,
            "beta": BETA,
            "seed": SEED,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": ("ising_sweep_x16f<0> (one sweep: colour 0 of each 16x16-row tile and its halos, then "
                       "colour 1)") if fused_sweep() else "ising_pass_x16u<3,0> (one colour pass)",
            "avg_launch_us": round(metro["launch_ms"] * 1e3, 3),
            "site_updates_per_ns_kernel": round(updates_per_launch / (metro["launch_ms"] * 1e6), 1),
        },
        # The launch is bound by VALU issue more than by bytes: instructions per
        # launch from the committed SQ counter summary, over the chip's measured
        # v_mul_lo_u32 issue rate, against this run's launch time.
        "valu_roof": valu_roof(metro["launch_ms"]),
        "metropolis": {k: (round(v, 6) if isinstance(v, float) else v) for k, v in metro.items()},
        "heat_bath": {k: (round(v, 6) if isinstance(v, float) else v) for k, v in hb.items()},
        "heat_bath_roofline_frac": round(bytes_per_launch / (hb["launch_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
        out["cpu_baseline_all_cores"] = cpu_baseline_all_cores(args.cpu_seconds / 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
