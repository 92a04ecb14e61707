"""Run configurations (CPU): the reference's schema tests, restated.

The cases follow the reference's own run-config tests
(ising/run_config.rs:281-729, potts/run_config.rs, gauge/run_config.rs and
config.rs): a known document parses to known fields, every config
round-trips through TOML, unknown keys and unrunnable values are refused at
load.  The hot start's stream (host code in the HIP library) is checked
against numpy's PCG64 from the seeded state and an independent restatement of
Canon's method.
"""
import os

import numpy as np
import pytest

from plaquette_amd.run import (ConfigError, GaugeRunConfig, IsingRunConfig, PottsRunConfig, check_dimension,
                               effective_schedule)

KNOWN = """
    shape = [16, 16]
    j = 1.0
    h = 0.25
    beta = 0.44
    updater = "metropolis"
    thermalize = 2000
    sweeps_between = 20
    n_samples = 1000
    seed = 12345
    start = "hot"
    description = "near T_c"
"""


def sample(cls=IsingRunConfig, **kw):
    base = dict(shape=[8, 8], j=1.0, beta=0.44, thermalize=100, sweeps_between=5, n_samples=200, seed=7)
    base.update(kw)
    return cls(**base)


def test_parses_a_known_toml_document():
    c = IsingRunConfig.parse(KNOWN)
    assert c.shape == [16, 16] and c.j == 1.0 and c.h == 0.25 and c.beta == 0.44
    assert (c.thermalize, c.sweeps_between, c.n_samples, c.seed) == (2000, 20, 1000, 12345)
    assert c.start == "hot" and c.updater == "metropolis" and c.description == "near T_c"
    assert c.schedule is None and c.backend == "cpu"


def test_optional_fields_default_when_omitted():
    c = IsingRunConfig.parse("shape = [4]\nj = 1.0\nbeta = 0.5\nthermalize = 0\nsweeps_between = 0\n"
                             "n_samples = 1\nseed = 1\n")
    assert (c.h, c.updater, c.schedule, c.backend, c.start, c.description) == (0.0, "metropolis", None, "cpu",
                                                                              "cold", None)
    assert effective_schedule(c.schedule) == "random"


@pytest.mark.parametrize("cls,extra", [
    (IsingRunConfig, {}),
    (IsingRunConfig, dict(schedule="checkerboard", updater="heat_bath", backend="gpu", start="hot")),
    (IsingRunConfig, dict(updater="swendsen_wang", backend="gpu")),
    (IsingRunConfig, dict(updater="wolff", description='a "quoted" note')),
    (PottsRunConfig, dict(h=[0.0, 0.5, -0.25])),
    (PottsRunConfig, dict(updater="swendsen_wang", backend="gpu", shape=[5, 7])),
    (GaugeRunConfig, dict(updater="heat_bath", schedule="checkerboard", backend="gpu", shape=[4, 4, 4, 4])),
])
def test_round_trips_through_toml(cls, extra):
    c = sample(cls, **extra)
    c.validate()
    text = c.to_toml()
    assert cls.parse(text) == c
    if "schedule" in extra:
        assert f'schedule = "{extra["schedule"]}"' in text


def test_description_is_omitted_from_dumps_when_absent():
    assert "description" not in sample().to_toml()
    assert "schedule" not in sample().to_toml()


def test_unknown_fields_are_rejected():
    text = KNOWN + 'updatr = "metropolis"\n'
    with pytest.raises(ConfigError) as e:
        IsingRunConfig.parse(text)
    assert e.value.kind == "parse"


def test_an_ising_file_is_refused_rather_than_run_as_a_potts_model():
    with pytest.raises(ConfigError) as e:
        PottsRunConfig.parse(KNOWN.replace('start = "hot"\n', ""))
    assert e.value.kind == "parse"


def test_missing_and_mistyped_fields_are_parse_errors():
    for text in [KNOWN.replace("seed = 12345\n", ""), KNOWN.replace("n_samples = 1000", "n_samples = -1"),
                 KNOWN.replace('updater = "metropolis"', 'updater = "glauber"'),
                 KNOWN.replace("shape = [16, 16]", 'shape = "16x16"'), "shape = [4\n"]:
        with pytest.raises(ConfigError) as e:
            IsingRunConfig.parse(text)
        assert e.value.kind == "parse"


@pytest.mark.parametrize("cls,extra,words", [
    (IsingRunConfig, dict(shape=[8, 0]), "positive"),
    (IsingRunConfig, dict(shape=[]), "at least 1"),
    (IsingRunConfig, dict(beta=0.0), "beta"),
    (IsingRunConfig, dict(beta=float("nan")), "beta"),
    (IsingRunConfig, dict(j=float("inf")), "j must be finite"),
    (IsingRunConfig, dict(h=float("nan")), "h must be finite"),
    (IsingRunConfig, dict(n_samples=0), "n_samples"),
    (IsingRunConfig, dict(backend="gpu"), "checkerboard"),
    (IsingRunConfig, dict(backend="gpu", schedule="random"), "checkerboard"),
    (IsingRunConfig, dict(backend="gpu", schedule="checkerboard", shape=[8, 7]), "odd on axis 1"),
    (IsingRunConfig, dict(updater="wolff", schedule="checkerboard"), "takes no schedule"),
    (IsingRunConfig, dict(updater="swendsen_wang", h=0.1), "spin-flip symmetry"),
    (PottsRunConfig, dict(h=[0.0, float("inf"), 0.0]), "entry 1"),
    (PottsRunConfig, dict(updater="swendsen_wang", h=[0.0, 0.1, 0.0]), "label symmetry"),
    (GaugeRunConfig, dict(shape=[8]), "at least 2"),
    (GaugeRunConfig, dict(updater="swendsen_wang"), "no pairwise bond graph"),
])
def test_validate_rejects_unrunnable_values(cls, extra, words):
    with pytest.raises(ConfigError) as e:
        sample(cls, **extra).validate()
    assert e.value.kind == "invalid"
    assert words in str(e.value)


def test_validate_accepts_what_the_reference_accepts():
    sample().validate()
    sample(thermalize=0, sweeps_between=0).validate()
    sample(updater="swendsen_wang", backend="gpu", shape=[7, 5]).validate()  # cluster: odd extents are fine
    sample(shape=[9, 3]).validate()                                          # CPU schedules: any shape
    sample(PottsRunConfig, updater="wolff", backend="gpu", shape=[3, 5]).validate()


def test_the_shape_length_is_the_dimension():
    c = sample(shape=[4, 4, 4])
    assert c.dimension() == 3
    c.check_dimension(3)
    with pytest.raises(ConfigError):
        c.check_dimension(2)
    with pytest.raises(ConfigError):
        check_dimension([4], 2)


def test_load_reads_a_file_from_disk(tmp_path):
    p = tmp_path / "run.toml"
    p.write_text(KNOWN)
    assert IsingRunConfig.load(p) == IsingRunConfig.parse(KNOWN)
    with pytest.raises(ConfigError) as e:
        IsingRunConfig.load(os.path.join(str(tmp_path), "missing.toml"))
    assert e.value.kind == "io"


# ------------------------------------------------------------- hot start ---

def _numpy_stream(seed):
    """numpy's PCG64 (same 128-bit multiplier and XSL-RR output) from the state
    rand's seed_from_u64 leaves; the LCG step is pinned in test_oracle.py."""
    import oracle as O

    g = O.pcg64(seed)
    bg = np.random.PCG64()
    st = bg.state
    st["state"]["state"] = g.state_lo | (g.state_hi << 64)
    st["state"]["inc"] = g.inc_lo | (g.inc_hi << 64)
    bg.state = st
    return bg


def _canon_below(bg, q):
    """Canon's method for a u32 range (the algorithm rand's random_range uses
    for small integer ranges): hi(x q) of a 32-bit draw, plus the carry a second
    draw decides when the low half is within q of overflowing."""
    x = int(bg.random_raw()) & 0xFFFFFFFF
    w = x * q
    result, lo = w >> 32, w & 0xFFFFFFFF
    if lo > (1 << 32) - q:
        x2 = int(bg.random_raw()) & 0xFFFFFFFF
        if lo + ((x2 * q) >> 32) > 0xFFFFFFFF:
            result += 1
    return result


@pytest.mark.parametrize("seed,q,n", [(20260728, 2, 4096), (7, 3, 1000), (12345, 16, 777), (0, 5, 10)])
def test_hot_start_draws_the_seeded_stream(seed, q, n):
    import plaquette_amd as P

    got, (state, inc) = P.hot_start(seed, q, n)
    bg = _numpy_stream(seed)
    want = [_canon_below(bg, q) for _ in range(n)]
    assert got.tolist() == want
    after = bg.state["state"]
    assert (state, inc) == (after["state"], after["inc"])


def test_hot_start_is_uniform_and_in_range():
    import plaquette_amd as P

    states, _ = P.hot_start(99, 3, 300000)
    counts = np.bincount(states, minlength=3)
    assert states.max() < 3
    assert np.all(np.abs(counts - 100000) < 5 * np.sqrt(300000 * (1 / 3) * (2 / 3)))
