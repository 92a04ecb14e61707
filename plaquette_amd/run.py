"""Run configurations and samplers: the caller side of the device chains.

A mirror of the reference's run plumbing for the device path:

- ``IsingRunConfig``, ``PottsRunConfig``, ``GaugeRunConfig``: the TOML schemas
  (ising/run_config.rs:52-117, potts/run_config.rs:56-120,
  gauge/run_config.rs:63-117), rejecting unknown keys, with the same defaults
  and the same ``validate`` rules (config.rs:176-255; ising/run_config.rs:168-205;
  potts/run_config.rs:221-262; gauge/run_config.rs:183-215).
- ``IsingSampler``, ``PottsSampler``, ``GaugeSampler``: build the chain in the
  reference's order -- seed the generator, draw a hot start from it, then the
  chain (ising/run_config.rs:237-251) -- thermalize once, and stream samples
  in batches of ``GPU_BATCH`` read back in one transfer each (device.rs:431-501;
  ising/sampler.rs:125-250).

Backends map onto the device chains as follows:

- ``backend = "gpu"``: the reference's device path.  Local rules run the
  checkerboard with the keyed draws under ``Profile.F32`` (the WGSL kernels'
  f32 arithmetic); Swendsen-Wang and Wolff run the cluster chain with the
  model's bond gap (2J for Ising, J for Potts).
- ``backend = "cpu"``: the reference's sequential CPU path.  The Ising
  checkerboard local rules run under ``Profile.CPU_STREAM``, continuing the
  run's own PCG64 stream after its start configuration, so the samples are the
  CPU run's.  The random schedule, CPU cluster updates and the other models'
  CPU streams have no device counterpart and are refused with a message.

The product has no CPU fallback: every sampler runs on the HIP library and
fails loudly without it.
"""
from __future__ import annotations

import collections
import math
from dataclasses import dataclass, field, fields
from typing import Iterator, List, Optional

import numpy as np

from . import (GpuChain, GpuClusterChain, Kernel, Model, Obs, Profile, Extent, Relabel, hot_start, local_desc)

GPU_BATCH = 64  # device.rs:497-501: samples per device round trip

RULES = ("metropolis", "heat_bath", "swendsen_wang", "wolff")
SCHEDULES = ("random", "checkerboard")
BACKENDS = ("cpu", "gpu")
STARTS = ("cold", "hot")


class ConfigError(ValueError):
    """config.rs:257-275: `kind` is "parse" (not valid TOML for the schema),
    "invalid" (parsed, but not runnable) or "io"."""

    _PREFIX = {"parse": "could not parse config", "invalid": "invalid config", "io": "could not read config file"}

    def __init__(self, kind: str, message: str):
        super().__init__(f"{self._PREFIX[kind]}: {message}")
        self.kind = kind


def _is_cluster(rule: str) -> bool:
    return rule in ("swendsen_wang", "wolff")  # config.rs:101-103


def effective_schedule(schedule: Optional[str]) -> str:
    """config.rs:167-171: an omitted schedule is the random one."""
    return schedule or "random"


def check_dimension(shape, expected: int) -> None:
    """config.rs:20-28."""
    if len(shape) != expected:
        raise ConfigError("invalid", f"this program is built for {expected} dimensions, "
                                     f"but shape{list(shape)} names {len(shape)}")


def check_shape(shape, min_dimension: int, scores: str) -> None:
    """config.rs:180-198."""
    if len(shape) < min_dimension:
        raise ConfigError("invalid", f"this run scores {scores}, which need at least {min_dimension} dimensions, "
                                     f"but shape{list(shape)} names {len(shape)}")
    for axis, l in enumerate(shape):
        if l == 0:
            raise ConfigError("invalid", f"every lattice extent must be positive, but shape{list(shape)} is 0 on "
                                         f"axis {axis}")


def check_updater(updater: str, schedule: Optional[str], backend: str, shape) -> None:
    """config.rs:214-244."""
    if _is_cluster(updater) and schedule is not None:
        raise ConfigError("invalid", f"{updater} builds clusters rather than walking variables, so it takes no "
                                     f"schedule; remove the schedule field")
    if backend == "gpu" and not _is_cluster(updater):
        if effective_schedule(schedule) != "checkerboard":
            raise ConfigError("invalid", f"the gpu backend runs {updater} as parallel color passes, which needs "
                                         f'schedule = "checkerboard"')
        for axis, l in enumerate(shape):
            if l % 2:
                raise ConfigError("invalid", f"the gpu checkerboard colors in parallel, which needs even extents, "
                                             f"but shape{list(shape)} is odd on axis {axis}")


def _check_beta(beta: float) -> None:
    if not math.isfinite(beta) or beta <= 0.0:
        raise ConfigError("invalid", f"beta must be positive and finite, got {beta}")


def _check_finite(name: str, v: float) -> None:
    if not math.isfinite(v):
        raise ConfigError("invalid", f"{name} must be finite, got {v}")


def _toml_value(v) -> str:
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        return repr(v) if math.isfinite(v) else ("nan" if v != v else ("inf" if v > 0 else "-inf"))
    if isinstance(v, str):
        out = []
        for ch in v:  # a TOML basic string: escape quotes, backslashes and control characters
            if ch in ('"', "\\"):
                out.append("\\" + ch)
            elif ord(ch) < 0x20 or ord(ch) == 0x7F:
                out.append("\\u%04x" % ord(ch))
            else:
                out.append(ch)
        return '"' + "".join(out) + '"'
    if isinstance(v, (list, tuple)):
        return "[" + ", ".join(_toml_value(x) for x in v) + "]"
    raise TypeError(f"cannot render {v!r}")


class _RunConfig:
    """Shared parse / load / to_toml over a dataclass schema."""

    _REQUIRED: tuple = ()
    _ENUMS = {"updater": RULES, "schedule": SCHEDULES, "backend": BACKENDS, "start": STARTS}
    _KINDS: dict = {}

    @classmethod
    def parse(cls, text: str):
        """Parse and validate TOML text (ising/run_config.rs:127-131)."""
        try:
            import tomli
        except ImportError as e:  # pragma: no cover - tomli ships with the image
            raise ConfigError("parse", f"no TOML parser available: {e}")
        try:
            raw = tomli.loads(text)
        except tomli.TOMLDecodeError as e:
            raise ConfigError("parse", str(e))
        names = {f.name for f in fields(cls)}
        unknown = sorted(set(raw) - names)
        if unknown:  # serde(deny_unknown_fields)
            raise ConfigError("parse", f"unknown field `{unknown[0]}`, expected one of {sorted(names)}")
        missing = [k for k in cls._REQUIRED if k not in raw]
        if missing:
            raise ConfigError("parse", f"missing field `{missing[0]}`")
        for key, value in raw.items():
            kind = cls._KINDS.get(key)
            if kind == "uint" and not (isinstance(value, int) and not isinstance(value, bool) and value >= 0):
                raise ConfigError("parse", f"invalid type for `{key}`: expected a non-negative integer")
            if kind == "float":
                if isinstance(value, bool) or not isinstance(value, (int, float)):
                    raise ConfigError("parse", f"invalid type for `{key}`: expected a float")
                raw[key] = float(value)
            if kind == "shape" and not (isinstance(value, list) and all(
                    isinstance(x, int) and not isinstance(x, bool) and x >= 0 for x in value)):
                raise ConfigError("parse", f"invalid type for `{key}`: expected a list of non-negative integers")
            if kind == "floats":
                if not (isinstance(value, list) and all(isinstance(x, (int, float)) and not isinstance(x, bool)
                                                        for x in value)):
                    raise ConfigError("parse", f"invalid type for `{key}`: expected a list of floats")
                raw[key] = [float(x) for x in value]
            if kind == "str" and not isinstance(value, str):
                raise ConfigError("parse", f"invalid type for `{key}`: expected a string")
            if key in cls._ENUMS and value not in cls._ENUMS[key]:
                raise ConfigError("parse", f"unknown variant `{value}` for `{key}`, expected one of "
                                           f"{list(cls._ENUMS[key])}")
        if "seed" in raw and raw["seed"] >= 1 << 64:
            raise ConfigError("parse", "invalid value for `seed`: it must fit in a u64")
        config = cls(**raw)
        config.validate()
        return config

    @classmethod
    def load(cls, path):
        """Read and validate a TOML file (ising/run_config.rs:120-125)."""
        try:
            with open(path, "r", encoding="utf-8") as f:
                text = f.read()
        except OSError as e:
            raise ConfigError("io", str(e)) from e
        return cls.parse(text)

    def dimension(self) -> int:
        return len(self.shape)

    def check_dimension(self, expected: int) -> None:
        check_dimension(self.shape, expected)

    def to_toml(self) -> str:
        """The config back as TOML; parsing it gives the same config (optional
        fields left unset are omitted, as serde's skip_serializing_if does)."""
        lines = []
        for f in fields(self):
            v = getattr(self, f.name)
            if v is None or (f.name == "h" and isinstance(v, list) and not v):
                continue
            lines.append(f"{f.name} = {_toml_value(v)}")
        return "\n".join(lines) + "\n"


@dataclass
class IsingRunConfig(_RunConfig):
    """ising/run_config.rs:52-117."""
    shape: List[int]
    j: float
    beta: float
    thermalize: int
    sweeps_between: int
    n_samples: int
    seed: int
    h: float = 0.0
    updater: str = "metropolis"
    schedule: Optional[str] = None
    backend: str = "cpu"
    start: str = "cold"
    description: Optional[str] = None

    _REQUIRED = ("shape", "j", "beta", "thermalize", "sweeps_between", "n_samples", "seed")
    _KINDS = {"shape": "shape", "j": "float", "beta": "float", "h": "float", "thermalize": "uint",
              "sweeps_between": "uint", "n_samples": "uint", "seed": "uint", "updater": "str", "schedule": "str",
              "backend": "str", "start": "str", "description": "str"}
    MIN_DIMENSION = 1

    def validate(self) -> None:
        """ising/run_config.rs:168-205."""
        check_shape(self.shape, self.MIN_DIMENSION, "nearest-neighbor bonds")
        _check_beta(self.beta)
        _check_finite("j", self.j)
        _check_finite("h", self.h)
        check_updater(self.updater, self.schedule, self.backend, self.shape)
        if _is_cluster(self.updater) and self.h != 0.0:
            raise ConfigError("invalid", f"{self.updater} relabels whole clusters at once, which needs the "
                                         f"spin-flip symmetry an external field breaks, but h is {self.h}")
        if self.n_samples == 0:
            raise ConfigError("invalid", "n_samples must be positive, or the run records nothing")


@dataclass
class PottsRunConfig(_RunConfig):
    """potts/run_config.rs:56-120; h is the per-label offset list (empty: symmetric)."""
    shape: List[int]
    j: float
    beta: float
    thermalize: int
    sweeps_between: int
    n_samples: int
    seed: int
    h: List[float] = field(default_factory=list)
    updater: str = "metropolis"
    schedule: Optional[str] = None
    backend: str = "cpu"
    start: str = "cold"
    description: Optional[str] = None

    _REQUIRED = ("shape", "j", "beta", "thermalize", "sweeps_between", "n_samples", "seed")
    _KINDS = {"shape": "shape", "j": "float", "beta": "float", "h": "floats", "thermalize": "uint",
              "sweeps_between": "uint", "n_samples": "uint", "seed": "uint", "updater": "str", "schedule": "str",
              "backend": "str", "start": "str", "description": "str"}
    MIN_DIMENSION = 1

    def validate(self) -> None:
        """potts/run_config.rs:221-262."""
        check_shape(self.shape, self.MIN_DIMENSION, "nearest-neighbor bonds")
        _check_beta(self.beta)
        _check_finite("j", self.j)
        for label, v in enumerate(self.h):
            if not math.isfinite(v):
                raise ConfigError("invalid", f"every entry of h must be finite, but entry {label} is {v}")
        check_updater(self.updater, self.schedule, self.backend, self.shape)
        if _is_cluster(self.updater) and any(v != 0.0 for v in self.h):
            raise ConfigError("invalid", f"{self.updater} relabels whole clusters at once, which needs the label "
                                         f"symmetry a per-label offset breaks, but h{self.h} is not all zero")
        if self.n_samples == 0:
            raise ConfigError("invalid", "n_samples must be positive, or the run records nothing")


@dataclass
class GaugeRunConfig(_RunConfig):
    """gauge/run_config.rs:63-117."""
    shape: List[int]
    j: float
    beta: float
    thermalize: int
    sweeps_between: int
    n_samples: int
    seed: int
    updater: str = "metropolis"
    schedule: Optional[str] = None
    backend: str = "cpu"
    start: str = "cold"
    description: Optional[str] = None

    _REQUIRED = ("shape", "j", "beta", "thermalize", "sweeps_between", "n_samples", "seed")
    _KINDS = {"shape": "shape", "j": "float", "beta": "float", "thermalize": "uint", "sweeps_between": "uint",
              "n_samples": "uint", "seed": "uint", "updater": "str", "schedule": "str", "backend": "str",
              "start": "str", "description": "str"}
    MIN_DIMENSION = 2

    def validate(self) -> None:
        """gauge/run_config.rs:183-215."""
        check_shape(self.shape, self.MIN_DIMENSION, "plaquettes")
        _check_beta(self.beta)
        _check_finite("j", self.j)
        if _is_cluster(self.updater):
            raise ConfigError("invalid", "the gauge action scores plaquettes, which have no pairwise bond graph "
                                         "for a cluster update to build on")
        check_updater(self.updater, self.schedule, self.backend, self.shape)
        if self.n_samples == 0:
            raise ConfigError("invalid", "n_samples must be positive, or the run records nothing")


class _Sampler:
    """Phasing shared by the three samplers (ising/sampler.rs:125-250)."""

    def __init__(self, chain, config, start: np.ndarray):
        self._chain = chain
        self.config = config
        self.shape = list(config.shape)
        self.start = start
        self._buffer: collections.deque = collections.deque()
        chain.advance(config.thermalize)

    @property
    def chain(self):
        """The device chain (GpuChain or GpuClusterChain) the sampler streams from."""
        return self._chain

    def samples(self) -> Iterator[np.ndarray]:
        """Thermalized configurations, one per sweeps_between sweeps, open-ended:
        bound it with itertools.islice.  Batches of GPU_BATCH are produced on the
        device and read back in one transfer; calling this again continues the
        same chain and buffer, as GpuChain::next_sample does (device.rs:490-495)."""
        while True:
            if not self._buffer:
                self._buffer.extend(self._chain.next_batch(GPU_BATCH))
            yield self._buffer.popleft()

    def observables(self) -> Iterator[Obs]:
        """The integer E/M (or plaquette) sums of each sample, measured on the
        device with no configuration read back.  Not interleaved with
        ``samples``: it advances the chain itself, sweeps_between per item."""
        if self._buffer:
            raise RuntimeError("observables() cannot continue a chain with buffered samples")
        while True:
            self._chain.advance(self.config.sweeps_between)
            yield self._chain.measure()

    def close(self) -> None:
        self._chain.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _start_states(config, q: int, n_vars: int):
    """The start configuration and the PCG64 stream state after it
    (ising/run_config.rs:237-251): cold is all state 0 and draws nothing."""
    if config.start == "cold":
        return np.zeros(n_vars, np.uint8), None
    return hot_start(config.seed, q, n_vars)


def _cluster_chain(config, q: int, gap: float, start: np.ndarray):
    extent, relabel = (Extent.ALL, Relabel.REDRAW) if config.updater == "swendsen_wang" else \
        (Extent.SEEDED, Relabel.FORCED_CHANGE)  # config.rs:128-140
    return GpuClusterChain.new(config.shape, q, gap, extent, relabel, config.beta, config.seed, start,
                               config.sweeps_between, Profile.F32)


def _kernel(rule: str) -> Kernel:
    return Kernel.METROPOLIS if rule == "metropolis" else Kernel.HEAT_BATH  # config.rs:112-118


class IsingSampler(_Sampler):
    """IsingSampler (ising/sampler.rs:103-250) over the device chains."""

    def __init__(self, config: IsingRunConfig):
        config.validate()
        n = int(np.prod(config.shape))
        start, stream = _start_states(config, 2, n)
        if config.backend == "gpu":
            if _is_cluster(config.updater):
                chain = _cluster_chain(config, 2, 2.0 * config.j, start)  # ising/mod.rs:152-154: gap 2J
            else:
                d = local_desc(config.shape, Model.ISING, 2, _kernel(config.updater), Profile.F32, config.beta,
                               config.j, config.h, None, config.seed, config.sweeps_between)
                chain = GpuChain.build(d, start)
        else:
            if _is_cluster(config.updater) or effective_schedule(config.schedule) != "checkerboard":
                raise ConfigError("invalid", "the device reproduces the CPU backend only for the checkerboard "
                                             "local rules (the CPU stream); the random schedule and the CPU "
                                             "cluster update run on the reference's CPU path")
            for axis, l in enumerate(config.shape):
                if l % 2:
                    raise ConfigError("invalid", f"the device checkerboard needs even extents, but shape"
                                                 f"{list(config.shape)} is odd on axis {axis}")
            d = local_desc(config.shape, Model.ISING, 2, _kernel(config.updater), Profile.CPU_STREAM,
                           config.beta, config.j, config.h, None, config.seed, config.sweeps_between)
            chain = GpuChain.build(d, start)
            if stream is not None:
                chain.set_pcg_state(*stream)  # the hot start's draws come first in the stream
        super().__init__(chain, config, start)


class PottsSampler(_Sampler):
    """PottsSampler (potts/sampler.rs) over the device chains; q is the
    compile-time Q of the reference's driver (POTTS_Q = 3 in its example)."""

    def __init__(self, config: PottsRunConfig, q: int = 3):
        config.validate()
        if config.h and len(config.h) != q:  # potts/run_config.rs:302-313
            raise ConfigError("invalid", f"this program is built for {q} labels, but h{config.h} gives "
                                         f"{len(config.h)}; leave h out for the symmetric model")
        if config.backend != "gpu":
            raise ConfigError("invalid", "the device reproduces the Potts model's gpu backend only; the CPU "
                                         "backend's stream runs on the reference's CPU path")
        n = int(np.prod(config.shape))
        start, _ = _start_states(config, q, n)
        if _is_cluster(config.updater):
            chain = _cluster_chain(config, q, config.j, start)  # potts/mod.rs:314-316: gap J
        else:
            d = local_desc(config.shape, Model.POTTS, q, _kernel(config.updater), Profile.F32, config.beta,
                           config.j, 0.0, config.h or None, config.seed, config.sweeps_between)
            chain = GpuChain.build(d, start)
        super().__init__(chain, config, start)
        self.q = q


class GaugeSampler(_Sampler):
    """GaugeSampler (gauge/sampler.rs:120-230) over the device link chain."""

    def __init__(self, config: GaugeRunConfig):
        config.validate()
        if config.backend != "gpu":
            raise ConfigError("invalid", "the device reproduces the gauge model's gpu backend only; the CPU "
                                         "backend's stream runs on the reference's CPU path")
        D = len(config.shape)
        n_links = int(np.prod(config.shape)) * D
        start, _ = _start_states(config, 2, n_links)
        d = local_desc(config.shape, Model.Z2, 2, _kernel(config.updater), Profile.F32, config.beta, config.j,
                       0.0, None, config.seed, config.sweeps_between)
        super().__init__(GpuChain.build(d, start), config, start)
