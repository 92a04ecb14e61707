"""plaquette_amd -- MI355X (gfx950) device path for plaquette's update seam.

Python mirror of the reference's device chain interface over the C ABI in
``include/plaquette_hip.h`` (``libplaquette_hip.so``):

* :class:`GpuChain` -- ``GpuChain`` (src/device.rs:590-701) built by the
  model constructors ``ising_gpu_chain`` (src/models/ising/gpu.rs:57-95),
  ``potts_gpu_chain`` (src/models/potts/gpu.rs) and ``gauge_gpu_chain``
  (src/models/gauge/gpu.rs:54-110);
* :class:`GpuClusterChain` -- ``GpuClusterChain`` (src/gpu_cluster.rs:144-637)
  with ``swendsen_wang`` / ``wolff`` constructors;
* :func:`gpu_available` -- ``Gpu::new() -> Option<Gpu>`` (src/device.rs:42-77).

Chains are iterators: ``next(chain)`` advances ``sweeps_between`` sweeps and
returns the configuration as a ``numpy.uint8`` array of state indices in
lattice order, like ``Iterator::next`` on the reference chains.  Errors the
reference raises as panics surface as :class:`PlaquetteError`.

The library is loaded from this directory and never replaced by a fallback:
if ``libplaquette_hip.so`` is missing, importing the chain types raises.
"""
from __future__ import annotations

import ctypes as C
import enum
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libplaquette_hip.so")

MAX_DIM = 8
MAX_Q = 16

PLQ_OK = 0
PLQ_EINVAL = -22
PLQ_ENODEV = -19
PLQ_ENOMEM = -12
PLQ_ENOCONV = -62
PLQ_EDEVICE = -5

_CODE_NAMES = {PLQ_EINVAL: "EINVAL", PLQ_ENODEV: "ENODEV", PLQ_ENOMEM: "ENOMEM", PLQ_ENOCONV: "ENOCONV",
               PLQ_EDEVICE: "EDEVICE"}


class Kernel(enum.IntEnum):
    """updater.rs Kernel: the single-variable rule a local chain applies."""
    METROPOLIS = 0
    HEAT_BATH = 1


class Model(enum.IntEnum):
    ISING = 0
    POTTS = 1
    Z2 = 2


class Profile(enum.IntEnum):
    """Acceptance arithmetic the integer thresholds reproduce."""
    F64 = 0   # the reference CPU path (updater.rs:77-167)
    F32 = 1   # the reference WGSL kernels (metropolis.wgsl, heat_bath.wgsl)
    CPU_STREAM = 2  # the CPU path's arithmetic and its own PCG64 stream (rng.rs:30-50), Ising only


class Extent(enum.IntEnum):
    """updater.rs Extent."""
    ALL = 0
    SEEDED = 1


class Relabel(enum.IntEnum):
    """updater.rs Relabel."""
    REDRAW = 0
    FORCED_CHANGE = 1


class PlaquetteError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{_CODE_NAMES.get(code, code)}: {message}")
        self.code = code
        self.message = message


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
        ("seed", C.c_uint64),
        ("sweeps_between", C.c_uint64),
    ]


class ClusterDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("shape", C.c_int64 * MAX_DIM),
        ("q", C.c_int32),
        ("extent", C.c_int32),
        ("relabel", C.c_int32),
        ("profile", C.c_int32),
        ("beta", C.c_double),
        ("gap", C.c_double),
        ("seed", C.c_uint64),
        ("sweeps_between", C.c_uint64),
    ]


class Obs(C.Structure):
    _fields_ = [
        ("sweep", C.c_uint64),
        ("n_vars", C.c_int64),
        ("bond_sum", C.c_int64),
        ("spin_sum", C.c_int64),
        ("counts", C.c_int64 * MAX_Q),
        ("energy", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {"sweep": self.sweep, "n_vars": self.n_vars, "bond_sum": self.bond_sum, "spin_sum": self.spin_sum,
                "counts": list(self.counts), "energy": self.energy}


class ClusterStats(C.Structure):
    _fields_ = [
        ("sweep", C.c_uint64),
        ("n_clusters", C.c_int64),
        ("largest", C.c_int64),
        ("moved", C.c_int64),
        ("size_hist", C.c_int64 * 40),
    ]


class EnsembleDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("n_replicas", C.c_int32),
        ("kernel", C.c_int32),
        ("profile", C.c_int32),
        ("shapes", C.POINTER(C.c_int64)),
        ("betas", C.POINTER(C.c_double)),
        ("seeds", C.POINTER(C.c_uint64)),
        ("j", C.c_double),
        ("measure_every", C.c_uint64),
        ("exchange", C.c_int32),
        ("exchange_seed", C.c_uint64),
        ("record_capacity", C.c_uint64),
    ]


# Every symbol the header declares, with its ctypes signature.
_P = C.POINTER
_HANDLE = C.c_void_p
SIGNATURES = {
    "plq_device_count": (C.c_int, [_P(C.c_int32)]),
    "plq_last_error": (C.c_char_p, []),
    "plq_chain_create_local": (C.c_int, [_P(LocalDesc), _P(C.c_uint8), C.c_size_t, _P(_HANDLE)]),
    "plq_chain_create_cluster": (C.c_int, [_P(ClusterDesc), _P(C.c_uint8), C.c_size_t, _P(_HANDLE)]),
    "plq_chain_advance": (C.c_int, [_HANDLE, C.c_uint64]),
    "plq_chain_next": (C.c_int, [_HANDLE, _P(C.c_uint8), C.c_size_t]),
    "plq_chain_read": (C.c_int, [_HANDLE, _P(C.c_uint8), C.c_size_t]),
    "plq_chain_next_batch": (C.c_int, [_HANDLE, C.c_uint64, _P(C.c_uint8), C.c_size_t]),
    "plq_hot_start": (C.c_int, [C.c_uint64, C.c_int32, _P(C.c_uint8), C.c_size_t, _P(C.c_uint64)]),
    "plq_chain_write": (C.c_int, [_HANDLE, _P(C.c_uint8), C.c_size_t]),
    "plq_chain_measure": (C.c_int, [_HANDLE, _P(Obs)]),
    "plq_chain_record": (C.c_int, [_HANDLE, C.c_uint64, C.c_uint64]),
    "plq_chain_series": (C.c_int, [_HANDLE, _P(Obs), C.c_uint64, _P(C.c_uint64)]),
    "plq_chain_sweeps_done": (C.c_int, [_HANDLE, _P(C.c_uint64)]),
    "plq_chain_set_sweep": (C.c_int, [_HANDLE, C.c_uint64]),
    "plq_chain_sync": (C.c_int, [_HANDLE]),
    "plq_chain_stream": (C.c_int, [_HANDLE, _P(C.c_void_p)]),
    "plq_chain_cluster_roots": (C.c_int, [_HANDLE, _P(C.c_int64), C.c_size_t]),
    "plq_chain_cluster_stats": (C.c_int, [_HANDLE, _P(ClusterStats)]),
    "plq_chain_set_pass_cap": (C.c_int, [_HANDLE, C.c_uint64]),
    "plq_chain_wolff_counts": (C.c_int, [_HANDLE, _P(C.c_uint64), _P(C.c_uint64)]),
    "plq_chain_record_clusters": (C.c_int, [_HANDLE, C.c_int]),
    "plq_chain_cluster_series": (C.c_int, [_HANDLE, C.c_void_p, C.c_uint64, _P(C.c_uint64)]),
    "plq_chain_destroy": (C.c_int, [_HANDLE]),
    "plq_chain_pcg_state": (C.c_int, [_HANDLE, _P(C.c_uint64)]),
    "plq_chain_set_pcg_state": (C.c_int, [_HANDLE, _P(C.c_uint64)]),
    "plq_chain_correlator": (C.c_int, [_HANDLE, C.c_int32, _P(C.c_int64), _P(C.c_double), C.c_size_t]),
    "plq_chain_correlator_noalloc": (C.c_int, [_HANDLE, C.c_int32, _P(C.c_int64), _P(C.c_double), C.c_size_t]),
    "plq_chain_polyakov_loop": (C.c_int, [_HANDLE, C.c_int32, _P(C.c_double)]),
    "plq_chain_wilson_rectangles": (C.c_int, [_HANDLE, C.c_int32, _P(C.c_double), C.c_size_t, _P(C.c_int32)]),
    "plq_ensemble_create": (C.c_int, [_P(EnsembleDesc), _P(C.c_uint8), C.c_size_t, _P(_HANDLE)]),
    "plq_ensemble_advance": (C.c_int, [_HANDLE, C.c_uint64]),
    "plq_ensemble_read": (C.c_int, [_HANDLE, _P(C.c_uint8), C.c_size_t]),
    "plq_ensemble_plaquette_sums": (C.c_int, [_HANDLE, _P(C.c_int64), C.c_size_t]),
    "plq_ensemble_betas": (C.c_int, [_HANDLE, _P(C.c_double), C.c_size_t]),
    "plq_ensemble_series": (C.c_int, [_HANDLE, _P(C.c_int64), _P(C.c_uint32), C.c_uint64, _P(C.c_uint64)]),
    "plq_ensemble_exchange_counts": (C.c_int, [_HANDLE, _P(C.c_uint64), _P(C.c_uint64), C.c_size_t,
                                               _P(C.c_size_t)]),
    "plq_ensemble_set_launch": (C.c_int, [_HANDLE, C.c_int32, C.c_int32, C.c_int32]),
    "plq_ensemble_tune": (C.c_int, [_HANDLE, _P(C.c_int32), _P(C.c_int32), _P(C.c_int32), _P(C.c_double)]),
    "plq_ensemble_sweeps_done": (C.c_int, [_HANDLE, _P(C.c_uint64)]),
    "plq_ensemble_stream": (C.c_int, [_HANDLE, _P(C.c_void_p)]),
    "plq_ensemble_sync": (C.c_int, [_HANDLE]),
    "plq_ensemble_destroy": (C.c_int, [_HANDLE]),
    "plq_blocknet_create": (C.c_int, [C.c_uint64, _P(C.c_void_p)]),
    "plq_blocknet_params": (C.c_int, [C.c_void_p, _P(C.c_float), C.c_size_t]),
    "plq_blocknet_set_params": (C.c_int, [C.c_void_p, _P(C.c_float), C.c_size_t]),
    "plq_blocknet_gradient": (C.c_int, [C.c_void_p, C.c_void_p, _P(C.c_float), C.c_size_t, _P(C.c_double)]),
    "plq_blocknet_train": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_double, _P(C.c_float)]),
    "plq_blocknet_logits": (C.c_int, [C.c_void_p, C.c_void_p, _P(C.c_float), C.c_size_t]),
    "plq_blocknet_destroy": (C.c_int, [C.c_void_p]),
    "plq_chain_block_sweeps": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "plq_chain_advance_with_blocks": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64]),
    "plq_chain_block_stats": (C.c_int, [C.c_void_p, _P(C.c_uint64), _P(C.c_uint64), _P(C.c_uint64)]),
    "plq_chain_block_logits": (C.c_int, [C.c_void_p, _P(C.c_float), C.c_size_t]),
    "plq_fold_seed": (C.c_uint32, [C.c_uint64]),
    "plq_threshold_table": (C.c_int, [_P(LocalDesc), _P(C.c_uint32), C.c_size_t]),
}

_lib = None


def lib():
    """Load ``libplaquette_hip.so``; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m plaquette_amd.build` "
                               "(there is no non-HIP fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc: int) -> int:
    if rc < 0:
        msg = lib().plq_last_error()
        raise PlaquetteError(rc, msg.decode() if msg else "")
    return rc


def gpu_available() -> bool:
    """Gpu::new().is_some() (device.rs:42-77)."""
    n = C.c_int32(0)
    return lib().plq_device_count(C.byref(n)) == PLQ_OK and n.value > 0


def hot_start(seed: int, q: int, n: int):
    """Configuration::hot from RandRng::seed_from_u64(seed) (configuration.rs:80-87,
    rng.rs:30-50): (states, stream state after the draws as (state, increment))."""
    out = np.empty(int(n), dtype=np.uint8)
    st = (C.c_uint64 * 4)()
    _check(lib().plq_hot_start(int(seed) & 0xFFFFFFFFFFFFFFFF, int(q), out.ctypes.data_as(_P(C.c_uint8)), out.size,
                               st))
    return out, (st[0] | (st[1] << 64), st[2] | (st[3] << 64))


def fold_seed(seed: int) -> int:
    return lib().plq_fold_seed(seed & 0xFFFFFFFFFFFFFFFF)


def _shape_arr(shape):
    arr = (C.c_int64 * MAX_DIM)()
    for i, l in enumerate(shape):
        arr[i] = int(l)
    return arr


def local_desc(shape, model=Model.ISING, q=2, kernel=Kernel.METROPOLIS, profile=Profile.F64, beta=0.44, j=1.0,
               h=0.0, offsets=None, seed=0, sweeps_between=1) -> LocalDesc:
    d = LocalDesc()
    d.dim = len(shape)
    d.shape = _shape_arr(shape)
    d.model = int(model)
    d.q = int(q)
    d.kernel = int(kernel)
    d.profile = int(profile)
    d.beta = float(beta)
    d.j = float(j)
    d.h = float(h)
    if offsets is not None:
        for i, o in enumerate(offsets):
            d.offsets[i] = float(o)
    d.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    d.sweeps_between = int(sweeps_between)
    return d


def threshold_table(desc: LocalDesc) -> np.ndarray:
    n = _check(lib().plq_threshold_table(C.byref(desc), None, 0))
    out = np.zeros(n, dtype=np.uint32)
    _check(lib().plq_threshold_table(C.byref(desc), out.ctypes.data_as(_P(C.c_uint32)), n))
    return out


class _Chain:
    """Common handle management for the two chain kinds."""

    def __init__(self, handle, n_vars: int, n_sites: int, sweeps_between: int):
        self._h = handle
        self.n_vars = n_vars
        self.n_sites = n_sites
        self.sweeps_between = sweeps_between
        self.shape = None   # extents, set by the constructors
        self.model = None   # Model of a local chain; None for a cluster chain

    # -- lifecycle ----------------------------------------------------------
    def close(self):
        if self._h:
            lib().plq_chain_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- the reference chain surface ----------------------------------------
    def advance(self, sweeps: int) -> None:
        """GpuChain::advance: enqueue sweeps without producing a snapshot."""
        _check(lib().plq_chain_advance(self._h, int(sweeps)))

    def __iter__(self):
        return self

    def __next__(self) -> np.ndarray:
        out = np.empty(self.n_vars, dtype=np.uint8)
        _check(lib().plq_chain_next(self._h, out.ctypes.data_as(_P(C.c_uint8)), out.size))
        return out

    def next_batch(self, count: int) -> np.ndarray:
        """`count` samples as GpuChain::run_batch yields them (device.rs:431-488):
        each after sweeps_between sweeps, read back in one transfer; shape
        (count, n_vars)."""
        out = np.empty((int(count), self.n_vars), dtype=np.uint8)
        _check(lib().plq_chain_next_batch(self._h, int(count), out.ctypes.data_as(_P(C.c_uint8)), out.size))
        return out

    def read(self) -> np.ndarray:
        out = np.empty(self.n_vars, dtype=np.uint8)
        _check(lib().plq_chain_read(self._h, out.ctypes.data_as(_P(C.c_uint8)), out.size))
        return out

    def write(self, states) -> None:
        st = np.ascontiguousarray(states, dtype=np.uint8)
        _check(lib().plq_chain_write(self._h, st.ctypes.data_as(_P(C.c_uint8)), st.size))

    def measure(self) -> Obs:
        o = Obs()
        _check(lib().plq_chain_measure(self._h, C.byref(o)))
        return o

    def record(self, every: int, capacity: int) -> None:
        _check(lib().plq_chain_record(self._h, int(every), int(capacity)))

    def series(self) -> list:
        n = C.c_uint64(0)
        _check(lib().plq_chain_series(self._h, None, 0, C.byref(n)))
        arr = (Obs * max(n.value, 1))()
        _check(lib().plq_chain_series(self._h, arr, n.value, C.byref(n)))
        return [arr[i] for i in range(n.value)]

    @property
    def sweeps_done(self) -> int:
        v = C.c_uint64(0)
        _check(lib().plq_chain_sweeps_done(self._h, C.byref(v)))
        return v.value

    def set_sweep(self, sweep: int) -> None:
        _check(lib().plq_chain_set_sweep(self._h, int(sweep)))

    def sync(self) -> None:
        _check(lib().plq_chain_sync(self._h))

    def stream(self) -> int:
        p = C.c_void_p(0)
        _check(lib().plq_chain_stream(self._h, C.byref(p)))
        return p.value or 0

    # -- per-configuration measurements on the device --------------------------
    def _correlator(self, model):
        if model is None:
            model = self.model
        if model is None:
            raise ValueError("a cluster chain needs the correlator convention: Model.ISING or Model.POTTS")
        lens = [int(l) // 2 + 1 for l in self.shape]
        sums = np.zeros(sum(lens), np.int64)
        vals = np.zeros(sum(lens), np.float64)
        _check(lib().plq_chain_correlator(self._h, int(model), sums.ctypes.data_as(_P(C.c_int64)),
                                          vals.ctypes.data_as(_P(C.c_double)), sums.size))
        cut = np.cumsum(lens)[:-1]
        return list(np.split(sums, cut)), list(np.split(vals, cut))

    def correlator(self, model=None) -> list:
        """Ising::correlator / Potts::correlator (ising/mod.rs:83-94,
        potts/mod.rs:218-233): one f64 array per axis, r = 0..L_mu/2."""
        return self._correlator(model)[1]

    def correlator_sums(self, model=None) -> list:
        """The raw pair sums behind the correlator (models/mod.rs:44-60)."""
        return self._correlator(model)[0]

    def polyakov_loop(self, direction: int) -> float:
        """Z2Gauge::polyakov_loop (gauge/mod.rs:178-201)."""
        v = C.c_double(0.0)
        _check(lib().plq_chain_polyakov_loop(self._h, int(direction), C.byref(v)))
        return v.value

    def wilson_rectangles(self, max_side: int) -> np.ndarray:
        """Z2Gauge::wilson_rectangles (gauge/mod.rs:224-268) as a (side+1) x (side+1) array."""
        side = C.c_int32(0)
        _check(lib().plq_chain_wilson_rectangles(self._h, int(max_side), None, 0, C.byref(side)))
        n = side.value + 1
        table = np.zeros(n * n, np.float64)
        _check(lib().plq_chain_wilson_rectangles(self._h, int(max_side), table.ctypes.data_as(_P(C.c_double)),
                                                 table.size, C.byref(side)))
        return table.reshape(n, n)


class GpuChain(_Chain):
    """GpuChain (device.rs:590-701): a checkerboard chain on one device."""

    # -- the CPU stream (Profile.CPU_STREAM) --

    def pcg_state(self):
        """(state, increment) of the chain's Lcg128Xsl64 stream as two ints."""
        v = (C.c_uint64 * 4)()
        _check(lib().plq_chain_pcg_state(self._h, v))
        return v[0] | (v[1] << 64), v[2] | (v[3] << 64)

    def set_pcg_state(self, state: int, increment: int) -> None:
        m = (1 << 64) - 1
        v = (C.c_uint64 * 4)(state & m, (state >> 64) & m, increment & m, (increment >> 64) & m)
        _check(lib().plq_chain_set_pcg_state(self._h, v))

    # -- learned block proposal (2D Ising, extents multiples of 16) --

    def block_sweeps(self, net: "BlockNet", sweeps: int) -> None:
        _check(lib().plq_chain_block_sweeps(self._h, net._h, int(sweeps)))

    def advance_with_blocks(self, net: "BlockNet", steps: int, plain_sweeps: int = 1) -> None:
        _check(lib().plq_chain_advance_with_blocks(self._h, net._h, int(steps), int(plain_sweeps)))

    def block_stats(self):
        """(accepted, proposed, block sweeps done)."""
        a, p, b = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        _check(lib().plq_chain_block_stats(self._h, C.byref(a), C.byref(p), C.byref(b)))
        return a.value, p.value, b.value

    def block_logits(self) -> np.ndarray:
        """Logits each block used in its latest pass, shape (n_blocks, 64)."""
        out = np.zeros(self.n_sites, np.float32)
        _check(lib().plq_chain_block_logits(self._h, out.ctypes.data_as(_P(C.c_float)), out.size))
        return out.reshape(-1, 64)

    @classmethod
    def build(cls, desc: LocalDesc, start) -> "GpuChain":
        st = np.ascontiguousarray(start, dtype=np.uint8).ravel()
        h = C.c_void_p(0)
        _check(lib().plq_chain_create_local(C.byref(desc), st.ctypes.data_as(_P(C.c_uint8)), st.size, C.byref(h)))
        shape = [int(desc.shape[i]) for i in range(desc.dim)]
        chain = cls(h, st.size, int(np.prod(shape)), desc.sweeps_between)
        chain.shape = shape
        chain.model = Model(desc.model)
        return chain


class BlockNet:
    """The learned block proposal's network (plq_blocknet_*): 32 boundary
    spins of an 8x8 block -> 64 logits; PLQ_BN_PARAMS FP32 parameters."""

    N_PARAMS = 9857

    def __init__(self, seed: int = 0):
        h = C.c_void_p(0)
        _check(lib().plq_blocknet_create(int(seed) & 0xFFFFFFFFFFFFFFFF, C.byref(h)))
        self._h = h

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(self.N_PARAMS, np.float32)
        _check(lib().plq_blocknet_params(self._h, out.ctypes.data_as(_P(C.c_float)), out.size))
        return out

    @params.setter
    def params(self, value) -> None:
        v = np.ascontiguousarray(value, dtype=np.float32).ravel()
        _check(lib().plq_blocknet_set_params(self._h, v.ctypes.data_as(_P(C.c_float)), v.size))

    def gradient(self, chain: GpuChain):
        """(mean loss, gradient) over every block of the chain's configuration."""
        g = np.zeros(self.N_PARAMS, np.float32)
        loss = C.c_double(0)
        _check(lib().plq_blocknet_gradient(self._h, chain._h, g.ctypes.data_as(_P(C.c_float)), g.size,
                                           C.byref(loss)))
        return loss.value, g

    def train(self, chain: GpuChain, steps: int, sweeps_between: int, lr: float) -> np.ndarray:
        losses = np.zeros(max(int(steps), 1), np.float32)
        _check(lib().plq_blocknet_train(self._h, chain._h, int(steps), int(sweeps_between), float(lr),
                                        losses.ctypes.data_as(_P(C.c_float))))
        return losses[: int(steps)]

    def logits(self, chain: GpuChain) -> np.ndarray:
        out = np.zeros(chain.n_sites, np.float32)
        _check(lib().plq_blocknet_logits(self._h, chain._h, out.ctypes.data_as(_P(C.c_float)), out.size))
        return out.reshape(-1, 64)

    def close(self) -> None:
        if self._h:
            lib().plq_blocknet_destroy(self._h)
            self._h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuClusterChain(_Chain):
    """GpuClusterChain (gpu_cluster.rs:144-637): SW / Wolff on one device."""

    @classmethod
    def new(cls, shape, q, gap, extent, relabel, beta, seed, start, sweeps_between=1,
            profile=Profile.F64) -> "GpuClusterChain":
        d = ClusterDesc()
        d.dim = len(shape)
        d.shape = _shape_arr(shape)
        d.q = int(q)
        d.extent = int(extent)
        d.relabel = int(relabel)
        d.profile = int(profile)
        d.beta = float(beta)
        d.gap = float(gap)
        d.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        d.sweeps_between = int(sweeps_between)
        st = np.ascontiguousarray(start, dtype=np.uint8).ravel()
        h = C.c_void_p(0)
        _check(lib().plq_chain_create_cluster(C.byref(d), st.ctypes.data_as(_P(C.c_uint8)), st.size, C.byref(h)))
        chain = cls(h, st.size, st.size, int(sweeps_between))
        chain.shape = [int(l) for l in shape]
        return chain

    @classmethod
    def swendsen_wang(cls, shape, q, gap, beta, seed, start, sweeps_between=1, profile=Profile.F64):
        return cls.new(shape, q, gap, Extent.ALL, Relabel.REDRAW, beta, seed, start, sweeps_between, profile)

    @classmethod
    def wolff(cls, shape, q, gap, beta, seed, start, sweeps_between=1, profile=Profile.F64):
        return cls.new(shape, q, gap, Extent.SEEDED, Relabel.FORCED_CHANGE, beta, seed, start, sweeps_between,
                       profile)

    def roots(self) -> np.ndarray:
        out = np.empty(self.n_sites, dtype=np.int64)
        _check(lib().plq_chain_cluster_roots(self._h, out.ctypes.data_as(_P(C.c_int64)), out.size))
        return out

    def stats(self) -> ClusterStats:
        s = ClusterStats()
        _check(lib().plq_chain_cluster_stats(self._h, C.byref(s)))
        return s

    def set_pass_cap(self, cap: int) -> None:
        _check(lib().plq_chain_set_pass_cap(self._h, int(cap)))

    def wolff_counts(self) -> tuple:
        """(sweeps that grew the seed's cluster, sweeps that ran the full
        decomposition) of a Wolff chain; (0, 0) while growth is off."""
        grown, full = C.c_uint64(), C.c_uint64()
        _check(lib().plq_chain_wolff_counts(self._h, C.byref(grown), C.byref(full)))
        return int(grown.value), int(full.value)

    def record_clusters(self, on: bool = True) -> None:
        """Keep each recorded sweep's cluster statistics on the device (with
        record(); read them with cluster_series())."""
        _check(lib().plq_chain_record_clusters(self._h, 1 if on else 0))

    def cluster_series(self) -> list:
        n = C.c_uint64()
        _check(lib().plq_chain_cluster_series(self._h, None, 0, C.byref(n)))
        if n.value == 0:
            return []
        buf = (ClusterStats * n.value)()
        _check(lib().plq_chain_cluster_series(self._h, C.cast(buf, C.c_void_p), n.value, C.byref(n)))
        return list(buf)


# -- model constructors (the reference's per-model gpu_chain functions) -------

def ising_gpu_chain(shape, j, h, beta, seed, start, sweeps_between=1, kernel=Kernel.METROPOLIS,
                    profile=Profile.F64) -> GpuChain:
    """models/ising/gpu.rs:57-95."""
    d = local_desc(shape, Model.ISING, 2, kernel, profile, beta, j, h, None, seed, sweeps_between)
    return GpuChain.build(d, start)


def potts_gpu_chain(shape, q, j, offsets, beta, seed, start, sweeps_between=1, kernel=Kernel.METROPOLIS,
                    profile=Profile.F64) -> GpuChain:
    """models/potts/gpu.rs (per-label offsets h[a])."""
    d = local_desc(shape, Model.POTTS, q, kernel, profile, beta, j, 0.0, offsets, seed, sweeps_between)
    return GpuChain.build(d, start)


def gauge_gpu_chain(shape, j, beta, seed, start, sweeps_between=1, kernel=Kernel.METROPOLIS,
                    profile=Profile.F64) -> GpuChain:
    """models/gauge/gpu.rs:54-110: links, link = site * D + dir."""
    d = local_desc(shape, Model.Z2, 2, kernel, profile, beta, j, 0.0, None, seed, sweeps_between)
    return GpuChain.build(d, start)


class Ensemble:
    """A batched Z2 gauge ensemble: many replicas of uneven shapes advanced
    together on one device, with optional replica exchange between
    neighbouring inverse temperatures of equal-shape replicas
    (plq_ensemble_* in include/plaquette_hip.h; no reference counterpart)."""

    def __init__(self, dim, shapes, betas, seeds, start=None, kernel=Kernel.HEAT_BATH, profile=Profile.F64,
                 j=1.0, measure_every=10, exchange=True, exchange_seed=0, record_capacity=4096):
        self.R = len(shapes)
        self.dim = dim
        self._shapes = np.ascontiguousarray(np.array(shapes, dtype=np.int64).reshape(self.R, dim))
        self._betas = np.ascontiguousarray(betas, dtype=np.float64)
        self._seeds = np.ascontiguousarray(np.array(seeds, dtype=np.uint64))
        self.sizes = [int(np.prod(s)) * dim for s in self._shapes]
        d = EnsembleDesc()
        d.dim = dim
        d.n_replicas = self.R
        d.kernel = int(kernel)
        d.profile = int(profile)
        d.shapes = self._shapes.ctypes.data_as(_P(C.c_int64))
        d.betas = self._betas.ctypes.data_as(_P(C.c_double))
        d.seeds = self._seeds.ctypes.data_as(_P(C.c_uint64))
        d.j = float(j)
        d.measure_every = int(measure_every)
        d.exchange = int(bool(exchange))
        d.exchange_seed = int(exchange_seed) & 0xFFFFFFFFFFFFFFFF
        d.record_capacity = int(record_capacity)
        st = np.zeros(sum(self.sizes), np.uint8) if start is None else np.ascontiguousarray(start, dtype=np.uint8)
        h = C.c_void_p(0)
        _check(lib().plq_ensemble_create(C.byref(d), st.ctypes.data_as(_P(C.c_uint8)), st.size, C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().plq_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def advance(self, sweeps: int) -> None:
        _check(lib().plq_ensemble_advance(self._h, int(sweeps)))

    def read(self) -> np.ndarray:
        out = np.empty(sum(self.sizes), np.uint8)
        _check(lib().plq_ensemble_read(self._h, out.ctypes.data_as(_P(C.c_uint8)), out.size))
        return out

    def plaquette_sums(self) -> np.ndarray:
        out = np.empty(self.R, np.int64)
        _check(lib().plq_ensemble_plaquette_sums(self._h, out.ctypes.data_as(_P(C.c_int64)), self.R))
        return out

    def betas(self) -> np.ndarray:
        out = np.empty(self.R, np.float64)
        _check(lib().plq_ensemble_betas(self._h, out.ctypes.data_as(_P(C.c_double)), self.R))
        return out

    def series(self):
        n = C.c_uint64(0)
        _check(lib().plq_ensemble_series(self._h, None, None, 0, C.byref(n)))
        sums = np.zeros(max(1, n.value) * self.R, np.int64)
        slots = np.zeros(max(1, n.value) * self.R, np.uint32)
        _check(lib().plq_ensemble_series(self._h, sums.ctypes.data_as(_P(C.c_int64)),
                                         slots.ctypes.data_as(_P(C.c_uint32)), n.value, C.byref(n)))
        k = n.value
        return sums[:k * self.R].reshape(k, self.R), slots[:k * self.R].reshape(k, self.R)

    def exchange_counts(self):
        n = C.c_size_t(0)
        _check(lib().plq_ensemble_exchange_counts(self._h, None, None, 0, C.byref(n)))
        acc = np.zeros(max(1, n.value), np.uint64)
        att = np.zeros(max(1, n.value), np.uint64)
        _check(lib().plq_ensemble_exchange_counts(self._h, acc.ctypes.data_as(_P(C.c_uint64)),
                                                  att.ctypes.data_as(_P(C.c_uint64)), n.value, C.byref(n)))
        return acc[:n.value], att[:n.value]

    def set_launch(self, block_size=256, per_lane=1, use_graph=True):
        """Lanes per block, links per lane, and the launch mode of a sweep block:
        use_graph 0 (or False) one launch per colour pass, 1 (or True) a captured
        hipGraph replay, 2 one cooperative launch with grid barriers."""
        _check(lib().plq_ensemble_set_launch(self._h, int(block_size), int(per_lane), int(use_graph)))

    def tune(self) -> dict:
        bs, pl, g = C.c_int32(0), C.c_int32(0), C.c_int32(0)
        us = C.c_double(0)
        _check(lib().plq_ensemble_tune(self._h, C.byref(bs), C.byref(pl), C.byref(g), C.byref(us)))
        return {"block_size": bs.value, "per_lane": pl.value, "use_graph": g.value,
                "us_per_sweep": us.value}

    @property
    def sweeps_done(self) -> int:
        v = C.c_uint64(0)
        _check(lib().plq_ensemble_sweeps_done(self._h, C.byref(v)))
        return v.value

    def stream(self) -> int:
        p = C.c_void_p(0)
        _check(lib().plq_ensemble_stream(self._h, C.byref(p)))
        return p.value or 0

    def sync(self) -> None:
        _check(lib().plq_ensemble_sync(self._h))
