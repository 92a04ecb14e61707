"""ctypes front of the CPU oracle (``oracle/libplq_oracle.so``).

TEST INFRASTRUCTURE ONLY.  This package is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import it.
The product (``plaquette_amd``) never imports it.

Each wrapper names the C function it calls; the C source cites the reference
lines (``src/updater.rs``, ``src/wgsl/*.wgsl``, ``src/lattice.rs``,
``src/cluster.rs``, ``src/statistics.rs``, ``src/models/*``) it restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libplq_oracle.so")

MODEL_ISING, MODEL_POTTS, MODEL_Z2 = 0, 1, 2
KERNEL_METROPOLIS, KERNEL_HEAT_BATH = 0, 1
PROFILE_F64, PROFILE_F32 = 0, 1
MAX_DIM, MAX_Q = 8, 16


def build() -> None:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "libplq_oracle.so"], check=True)


class LocalDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("shape", C.c_int64 * MAX_DIM),
        ("model", C.c_int32),
        ("q", C.c_int32),
        ("kernel", C.c_int32),
        ("profile", C.c_int32),
        ("beta", C.c_double),
        ("j", C.c_double),
        ("h", C.c_double),
        ("offsets", C.c_double * MAX_Q),
        ("seed32", C.c_uint32),
    ]


class ClusterDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("shape", C.c_int64 * MAX_DIM),
        ("q", C.c_int32),
        ("seeded", C.c_int32),
        ("forced", C.c_int32),
        ("profile", C.c_int32),
        ("beta", C.c_double),
        ("gap", C.c_double),
        ("seed32", C.c_uint32),
    ]


class Estimate(C.Structure):
    _fields_ = [("mean", C.c_double), ("stderr", C.c_double), ("tau_int", C.c_double), ("n_eff", C.c_double)]


class Derived(C.Structure):
    _fields_ = [("value", C.c_double), ("stderr", C.c_double), ("n_blocks", C.c_int64)]


class EnsembleDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("n_replicas", C.c_int32),
        ("kernel", C.c_int32),
        ("profile", C.c_int32),
        ("shapes", C.POINTER(C.c_int64)),
        ("betas", C.POINTER(C.c_double)),
        ("seed32s", C.POINTER(C.c_uint32)),
        ("j", C.c_double),
        ("measure_every", C.c_uint64),
        ("exchange", C.c_int32),
        ("ex_seed32", C.c_uint32),
    ]


class Pcg64(C.Structure):
    _fields_ = [("state_lo", C.c_uint64), ("state_hi", C.c_uint64), ("inc_lo", C.c_uint64), ("inc_hi", C.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER
        sig = {
            "orc_lowbias32": (C.c_uint32, [C.c_uint32]),
            "orc_keyed_bits": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
            "orc_keyed_uniform": (C.c_float, [C.c_uint32, C.c_uint32, C.c_uint32]),
            "orc_fold_seed": (C.c_uint32, [C.c_uint64]),
            "orc_n_sites": (C.c_int64, [C.c_int32, P(C.c_int64)]),
            "orc_neighbor_table": (None, [C.c_int32, P(C.c_int64), P(C.c_int64)]),
            "orc_parity_table": (None, [C.c_int32, P(C.c_int64), P(C.c_uint8)]),
            "orc_plaquette_links": (None, [C.c_int32, P(C.c_int64), C.c_int64, P(C.c_int64)]),
            "orc_staple_table": (None, [C.c_int32, P(C.c_int64), P(C.c_int64)]),
            "orc_local_sweeps": (C.c_int, [P(LocalDesc), P(C.c_uint8), C.c_uint32, C.c_uint32]),
            "orc_max_threads": (C.c_int32, []),
            "orc_local_sweeps_mt": (C.c_double, [P(LocalDesc), P(C.c_uint8), C.c_uint32, C.c_uint32, C.c_int32]),
            "orc_ising_metropolis_accept": (C.c_int, [P(LocalDesc), C.c_int32, C.c_int64, C.c_uint32]),
            "orc_ising_heat_bath_choice": (C.c_int, [P(LocalDesc), C.c_int32, C.c_int64, C.c_uint32]),
            "orc_cluster_sweeps": (C.c_int, [P(ClusterDesc), P(C.c_uint8), C.c_uint32, C.c_uint32, P(C.c_int64)]),
            "orc_cluster_bonds": (C.c_int, [P(ClusterDesc), P(C.c_uint8), C.c_uint32, P(C.c_uint8)]),
            "orc_components": (C.c_int, [C.c_int32, P(C.c_int64), P(C.c_uint8), P(C.c_int64)]),
            "orc_ising_sums": (None, [C.c_int32, P(C.c_int64), P(C.c_uint8), P(C.c_int64), P(C.c_int64)]),
            "orc_potts_sums": (None, [C.c_int32, P(C.c_int64), C.c_int32, P(C.c_uint8), P(C.c_int64), P(C.c_int64)]),
            "orc_plaquette_sum": (C.c_int64, [C.c_int32, P(C.c_int64), P(C.c_uint8)]),
            "orc_reduce": (Estimate, [P(C.c_double), C.c_int64]),
            "orc_specific_heat": (Derived, [P(C.c_double), C.c_int64, C.c_double, C.c_double]),
            "orc_susceptibility": (Derived, [P(C.c_double), C.c_int64, C.c_double, C.c_double]),
            "orc_binder": (Derived, [P(C.c_double), C.c_int64]),
            "orc_pcg64_seed_from_u64": (None, [P(Pcg64), C.c_uint64]),
            "orc_pcg64_next_u64": (C.c_uint64, [P(Pcg64)]),
            "orc_pcg64_next_f64": (C.c_double, [P(Pcg64)]),
            "orc_ensemble_init": (C.c_int, [P(EnsembleDesc), P(C.c_uint32), P(C.c_uint32)]),
            "orc_ensemble_sweeps": (C.c_int, [P(EnsembleDesc), P(C.c_uint8), P(C.c_uint32), P(C.c_uint32),
                                              P(C.c_uint64), P(C.c_uint64), C.c_uint64, P(C.c_int64),
                                              P(C.c_uint32), C.c_uint64, P(C.c_uint64), P(C.c_uint64)]),
            "orc_expf": (C.c_float, [C.c_float]),
            "orc_bf16": (C.c_float, [C.c_float]),
            "orc_blocknet_logits": (None, [P(C.c_float), C.c_int, C.c_int64, C.c_int64, P(C.c_uint8),
                                           P(C.c_double)]),
            "orc_blocknet_loss_grad": (C.c_double, [P(C.c_float), C.c_int, C.c_int64, C.c_int64, P(C.c_uint8),
                                                    P(C.c_double)]),
            "orc_block_mh_pass": (C.c_int64, [C.c_int64, C.c_int64, P(C.c_uint8), P(C.c_float), C.c_int,
                                              C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double]),
            "orc_pmode_ising_sweeps": (C.c_double, [C.c_int32, P(C.c_int64), C.c_int32, C.c_double, C.c_double,
                                                    C.c_double, C.c_uint64, P(C.c_uint8), C.c_uint32]),
            "orc_pmode_ising_sweeps_from": (C.c_double, [C.c_int32, P(C.c_int64), C.c_int32, C.c_double,
                                                         C.c_double, C.c_double, P(Pcg64), P(C.c_uint8),
                                                         C.c_uint32]),
            "orc_axis_pair_sums": (None, [C.c_int32, P(C.c_int64), P(C.c_uint8), C.c_int32, P(C.c_int64)]),
            "orc_polyakov_sum": (C.c_int64, [C.c_int32, P(C.c_int64), P(C.c_uint8), C.c_int32, P(C.c_int64)]),
            "orc_wilson_sums": (C.c_int32, [C.c_int32, P(C.c_int64), P(C.c_uint8), C.c_int32, P(C.c_int64),
                                            P(C.c_int64)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _shape(shape):
    arr = (C.c_int64 * MAX_DIM)()
    for i, l in enumerate(shape):
        arr[i] = int(l)
    return arr


def _u8(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def _i64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def _f64(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ------------------------------------------------------------------ RNG ---

def lowbias32(x: int) -> int:
    return lib().orc_lowbias32(x & 0xFFFFFFFF)


def keyed_bits(seed32: int, var: int, sweep: int) -> int:
    return lib().orc_keyed_bits(seed32 & 0xFFFFFFFF, var & 0xFFFFFFFF, sweep & 0xFFFFFFFF)


def keyed_uniform(seed32: int, var: int, sweep: int) -> float:
    return lib().orc_keyed_uniform(seed32 & 0xFFFFFFFF, var & 0xFFFFFFFF, sweep & 0xFFFFFFFF)


def fold_seed(seed: int) -> int:
    return lib().orc_fold_seed(seed & 0xFFFFFFFFFFFFFFFF)


# ------------------------------------------------------------- geometry ---

def n_sites(shape) -> int:
    return int(np.prod(shape))


def neighbor_table(shape) -> np.ndarray:
    d = len(shape)
    out = np.zeros((n_sites(shape), 2 * d), dtype=np.int64)
    lib().orc_neighbor_table(d, _shape(shape), _i64(out))
    return out


def parity_table(shape) -> np.ndarray:
    out = np.zeros(n_sites(shape), dtype=np.uint8)
    lib().orc_parity_table(len(shape), _shape(shape), _u8(out))
    return out


def plaquette_links(shape, p: int) -> list:
    out = np.zeros(4, dtype=np.int64)
    lib().orc_plaquette_links(len(shape), _shape(shape), p, _i64(out))
    return [int(x) for x in out]


def staple_table(shape) -> np.ndarray:
    d = len(shape)
    out = np.zeros((n_sites(shape) * d, 6 * (d - 1)), dtype=np.int64)
    lib().orc_staple_table(d, _shape(shape), _i64(out))
    return out


# --------------------------------------------------------------- sweeps ---

def local_desc(shape, model=MODEL_ISING, q=2, kernel=KERNEL_METROPOLIS, profile=PROFILE_F64,
               beta=0.44, j=1.0, h=0.0, offsets=None, seed=0) -> LocalDesc:
    d = LocalDesc()
    d.dim = len(shape)
    d.shape = _shape(shape)
    d.model = model
    d.q = q
    d.kernel = kernel
    d.profile = profile
    d.beta = beta
    d.j = j
    d.h = h
    if offsets is not None:
        for i, o in enumerate(offsets):
            d.offsets[i] = float(o)
    d.seed32 = fold_seed(seed)
    return d


def local_sweeps(desc: LocalDesc, state: np.ndarray, sweep0: int, n_sweeps: int) -> np.ndarray:
    st = np.ascontiguousarray(state, dtype=np.uint8).copy()
    rc = lib().orc_local_sweeps(C.byref(desc), _u8(st), sweep0, n_sweeps)
    if rc != 0:
        raise ValueError("orc_local_sweeps rejected the description")
    return st


def max_threads() -> int:
    """Threads an OpenMP region of the oracle uses (OMP_NUM_THREADS, else the CPUs)."""
    return int(lib().orc_max_threads())


def local_sweeps_mt(desc: LocalDesc, state: np.ndarray, sweep0: int, n_sweeps: int, threads: int = 0):
    """(state, seconds): the same sweeps over OpenMP threads (0 = all)."""
    st = np.ascontiguousarray(state, dtype=np.uint8).copy()
    secs = lib().orc_local_sweeps_mt(C.byref(desc), _u8(st), sweep0, n_sweeps, threads)
    if secs < 0:
        raise ValueError("bad descriptor")
    return st, secs


def ising_metropolis_accept(desc: LocalDesc, cur: int, nsum: int, m: int) -> int:
    return lib().orc_ising_metropolis_accept(C.byref(desc), cur, nsum, m)


def ising_heat_bath_choice(desc: LocalDesc, cur: int, nsum: int, m: int) -> int:
    return lib().orc_ising_heat_bath_choice(C.byref(desc), cur, nsum, m)


def cluster_desc(shape, q=3, seeded=False, forced=False, profile=PROFILE_F64, beta=1.0, gap=1.0,
                 seed=0) -> ClusterDesc:
    d = ClusterDesc()
    d.dim = len(shape)
    d.shape = _shape(shape)
    d.q = q
    d.seeded = int(seeded)
    d.forced = int(forced)
    d.profile = profile
    d.beta = beta
    d.gap = gap
    d.seed32 = fold_seed(seed)
    return d


def cluster_sweeps(desc: ClusterDesc, state: np.ndarray, sweep0: int, n_sweeps: int):
    st = np.ascontiguousarray(state, dtype=np.uint8).copy()
    roots = np.zeros(st.size, dtype=np.int64)
    rc = lib().orc_cluster_sweeps(C.byref(desc), _u8(st), sweep0, n_sweeps, _i64(roots))
    if rc != 0:
        raise ValueError("orc_cluster_sweeps rejected the description")
    return st, roots


def cluster_bonds(desc: ClusterDesc, state: np.ndarray, sweep: int) -> np.ndarray:
    st = np.ascontiguousarray(state, dtype=np.uint8)
    bonds = np.zeros(st.size * desc.dim, dtype=np.uint8)
    lib().orc_cluster_bonds(C.byref(desc), _u8(st), sweep, _u8(bonds))
    return bonds


def components(shape, bonds: np.ndarray) -> np.ndarray:
    b = np.ascontiguousarray(bonds, dtype=np.uint8)
    roots = np.zeros(n_sites(shape), dtype=np.int64)
    lib().orc_components(len(shape), _shape(shape), _u8(b), _i64(roots))
    return roots


# ---------------------------------------------------------- observables ---

def ising_sums(shape, state):
    st = np.ascontiguousarray(state, dtype=np.uint8)
    b, s = C.c_int64(), C.c_int64()
    lib().orc_ising_sums(len(shape), _shape(shape), _u8(st), C.byref(b), C.byref(s))
    return b.value, s.value


def potts_sums(shape, q, state):
    st = np.ascontiguousarray(state, dtype=np.uint8)
    a = C.c_int64()
    counts = np.zeros(q, dtype=np.int64)
    lib().orc_potts_sums(len(shape), _shape(shape), q, _u8(st), C.byref(a), _i64(counts))
    return a.value, counts


def plaquette_sum(shape, links) -> int:
    st = np.ascontiguousarray(links, dtype=np.uint8)
    return lib().orc_plaquette_sum(len(shape), _shape(shape), _u8(st))


# ------------------------------------------------------- measurements ---

def axis_pair_sums(shape, state, pair_kind: int) -> list:
    """Raw pair sums per axis (models/mod.rs:44-60); pair_kind 0 Ising, 1 Potts."""
    st = np.ascontiguousarray(state, dtype=np.uint8)
    lens = [int(l) // 2 + 1 for l in shape]
    out = np.zeros(sum(lens), dtype=np.int64)
    lib().orc_axis_pair_sums(len(shape), _shape(shape), _u8(st), pair_kind, _i64(out))
    return list(np.split(out, np.cumsum(lens)[:-1]))


def ising_correlator(shape, state) -> list:
    """Ising::correlator (ising/mod.rs:83-94): sum / N per axis and displacement."""
    n = float(np.prod(shape))
    return [s.astype(np.float64) / n for s in axis_pair_sums(shape, state, 0)]


def potts_correlator(shape, q, state) -> list:
    """Potts::correlator (potts/mod.rs:218-233): agreements / N - 1/q."""
    n = float(np.prod(shape))
    return [s.astype(np.float64) / n - 1.0 / q for s in axis_pair_sums(shape, state, 1)]


def polyakov_loop(shape, links, direction: int) -> float:
    """Z2Gauge::polyakov_loop (gauge/mod.rs:178-201)."""
    st = np.ascontiguousarray(links, dtype=np.uint8)
    lines = C.c_int64()
    total = lib().orc_polyakov_sum(len(shape), _shape(shape), _u8(st), direction, C.byref(lines))
    return float(total) / float(lines.value)


def wilson_sums(shape, links, max_side: int):
    """(totals, placements) of Z2Gauge::wilson_rectangles (gauge/mod.rs:224-268)."""
    st = np.ascontiguousarray(links, dtype=np.uint8)
    shortest = min(int(l) for l in shape)
    side = min(max_side, shortest // 2)
    totals = np.zeros((side + 1) * (side + 1), dtype=np.int64)
    placements = C.c_int64()
    got = lib().orc_wilson_sums(len(shape), _shape(shape), _u8(st), max_side, _i64(totals), C.byref(placements))
    assert got == side
    return totals.reshape(side + 1, side + 1), placements.value


def wilson_rectangles(shape, links, max_side: int) -> np.ndarray:
    """The table itself: 1.0 on row and column 0, totals / placements elsewhere."""
    totals, placements = wilson_sums(shape, links, max_side)
    table = np.ones(totals.shape, dtype=np.float64)
    table[1:, 1:] = totals[1:, 1:].astype(np.float64) / float(placements)
    return table


# ----------------------------------------------------------- statistics ---

def reduce(series):
    x = np.ascontiguousarray(series, dtype=np.float64)
    return lib().orc_reduce(_f64(x), x.size)


def specific_heat(e, beta, n):
    x = np.ascontiguousarray(e, dtype=np.float64)
    return lib().orc_specific_heat(_f64(x), x.size, beta, n)


def susceptibility(m, beta, n):
    x = np.ascontiguousarray(m, dtype=np.float64)
    return lib().orc_susceptibility(_f64(x), x.size, beta, n)


def binder(m):
    x = np.ascontiguousarray(m, dtype=np.float64)
    return lib().orc_binder(_f64(x), x.size)


# ------------------------------------------------------------- P-mode ---

def pcg64(seed: int) -> Pcg64:
    g = Pcg64()
    lib().orc_pcg64_seed_from_u64(C.byref(g), seed)
    return g


def pmode_ising_sweeps_from(shape, kernel, beta, j, h, gen: Pcg64, state, n_sweeps):
    """The CPU loop continuing `gen` (advanced in place)."""
    st = np.ascontiguousarray(state, dtype=np.uint8).copy()
    lib().orc_pmode_ising_sweeps_from(len(shape), _shape(shape), kernel, beta, j, h, C.byref(gen), _u8(st), n_sweeps)
    return st


def pmode_ising_sweeps(shape, kernel, beta, j, h, seed, state, n_sweeps):
    st = np.ascontiguousarray(state, dtype=np.uint8).copy()
    secs = lib().orc_pmode_ising_sweeps(len(shape), _shape(shape), kernel, beta, j, h, seed, _u8(st), n_sweeps)
    return st, secs


# ------------------------------------------------------------- ensemble ---

EXCHANGE_SALT = 0x68E31DA4


class Ensemble:
    """CPU restatement of a batched Z2 ensemble with replica exchange."""

    def __init__(self, dim, shapes, betas, seeds, kernel=KERNEL_HEAT_BATH, profile=PROFILE_F64, j=1.0,
                 measure_every=10, exchange=True, exchange_seed=0, state=None):
        self.R = len(shapes)
        self._shapes = np.ascontiguousarray(np.array(shapes, dtype=np.int64).reshape(self.R, dim))
        self._betas = np.ascontiguousarray(betas, dtype=np.float64)
        self._seeds = np.array([fold_seed(s) for s in seeds], dtype=np.uint32)
        d = EnsembleDesc()
        d.dim = dim
        d.n_replicas = self.R
        d.kernel = kernel
        d.profile = profile
        d.shapes = self._shapes.ctypes.data_as(C.POINTER(C.c_int64))
        d.betas = self._betas.ctypes.data_as(C.POINTER(C.c_double))
        d.seed32s = self._seeds.ctypes.data_as(C.POINTER(C.c_uint32))
        d.j = j
        d.measure_every = measure_every
        d.exchange = int(exchange)
        d.ex_seed32 = fold_seed(exchange_seed) ^ EXCHANGE_SALT
        self.desc = d
        self.dim = dim
        self.sizes = [int(np.prod(s)) * dim for s in self._shapes]
        self.state = np.ascontiguousarray(state, dtype=np.uint8).copy() if state is not None else \
            np.zeros(sum(self.sizes), np.uint8)
        self.slot_of = np.zeros(self.R, np.uint32)
        self.rep_at = np.zeros(self.R, np.uint32)
        lib().orc_ensemble_init(C.byref(d), self.slot_of.ctypes.data_as(C.POINTER(C.c_uint32)),
                                self.rep_at.ctypes.data_as(C.POINTER(C.c_uint32)))
        self.sweeps_done = C.c_uint64(0)
        self.rounds_done = C.c_uint64(0)
        self.records = []
        self.slots = []
        self.n_pairs = sum(max(0, c - 1) for c in self._group_sizes())
        self.accepted = np.zeros(max(1, self.n_pairs), np.uint64)

    def _group_sizes(self):
        seen = []
        counts = []
        for s in self._shapes.tolist():
            if s in seen:
                counts[seen.index(s)] += 1
            else:
                seen.append(s)
                counts.append(1)
        return counts

    def advance(self, sweeps):
        cap = sweeps + 1
        rec = np.zeros(cap * self.R, np.int64)
        slots = np.zeros(cap * self.R, np.uint32)
        n = C.c_uint64(0)
        P = C.POINTER
        lib().orc_ensemble_sweeps(C.byref(self.desc), _u8(self.state), self.slot_of.ctypes.data_as(P(C.c_uint32)),
                                  self.rep_at.ctypes.data_as(P(C.c_uint32)), C.byref(self.sweeps_done),
                                  C.byref(self.rounds_done), sweeps, rec.ctypes.data_as(P(C.c_int64)),
                                  slots.ctypes.data_as(P(C.c_uint32)), cap, C.byref(n),
                                  self.accepted.ctypes.data_as(P(C.c_uint64)))
        for k in range(n.value):
            self.records.append(rec[k * self.R:(k + 1) * self.R].copy())
            self.slots.append(slots[k * self.R:(k + 1) * self.R].copy())
        return self

    def betas_now(self):
        return self._betas[self.slot_of]


# -- learned block proposal (plq_oracle_block.c) --------------------------------

BN_PARAMS = 9857
BN_CHANNELS = 32
BLOCK_PROPOSAL_KEY = 0x3C6EF372
BLOCK_ACCEPT_KEY = 0xA54FF53A


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


def expf(x: float) -> float:
    """orc_expf: the deterministic f32 exponential of the block update."""
    return float(lib().orc_expf(float(x)))


def bf16(x: float) -> float:
    return float(lib().orc_bf16(float(x)))


def blocknet_logits(params, shape, state, round_bf16=True) -> np.ndarray:
    """orc_blocknet_logits: (n_blocks, 64) logits, FP64."""
    L0, L1 = shape
    p, pp = _f32(params)
    assert p.size == BN_PARAMS
    st = np.ascontiguousarray(state, dtype=np.uint8)
    out = np.zeros((L0 // 8) * (L1 // 8) * 64, np.float64)
    lib().orc_blocknet_logits(pp, int(round_bf16), L0, L1, _u8(st),
                              out.ctypes.data_as(C.POINTER(C.c_double)))
    return out.reshape(-1, 64)


def blocknet_loss_grad(params, shape, state, round_bf16=True):
    """orc_blocknet_loss_grad: (mean BCE loss, gradient[BN_PARAMS])."""
    L0, L1 = shape
    p, pp = _f32(params)
    st = np.ascontiguousarray(state, dtype=np.uint8)
    g = np.zeros(BN_PARAMS, np.float64)
    loss = lib().orc_blocknet_loss_grad(pp, int(round_bf16), L0, L1, _u8(st),
                                        g.ctypes.data_as(C.POINTER(C.c_double)))
    return loss, g


def block_mh_pass(shape, state, logits, colour, seed32, bstep, beta, j, h):
    """orc_block_mh_pass on a copy of state; returns (new state, accepted)."""
    L0, L1 = shape
    st = np.array(state, dtype=np.uint8, copy=True)
    lg, lp = _f32(np.asarray(logits).reshape(-1))
    acc = lib().orc_block_mh_pass(L0, L1, _u8(st), lp, int(colour),
                                  int(seed32) & 0xFFFFFFFF, int(bstep) & 0xFFFFFFFF, float(beta), float(j), float(h))
    return st, int(acc)
