"""Randomized bit-exact parity over the device paths (GPU): fixed seeds draw
lattices (1D to 4D, powers of two and other even extents, up to 300,000
sites), models (Ising with h and antiferromagnetic J, Potts q = 3..6 with
label offsets, Z2 links, Swendsen-Wang and Wolff with either relabel rule),
kernels, threshold profiles, couplings from beta = 0.05 to 4, hot and cold
starts, and the launch switches that select other kernels (the uniform
pass, the one-workgroup chains, the pass grid cap, Wolff growth and its
level budget).  Each case runs a few sweeps on the device and on the oracle
and compares the states after every sweep, and the cluster partition after
the last.  The cases complement the hand-picked ones elsewhere: the 4D
class collision of the uniform pass was a combination none of them held."""
import os
import random

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

SWITCHES = ["PLQ_X16U", "PLQ_SMALL_CHAIN", "PLQ_PASS_BLOCKS", "PLQ_SMALL_CLUSTER", "PLQ_WOLFF_GROW",
            "PLQ_WOLFF_LEVELS"]


def _shape(rng, D, maxn, even):
    while True:
        if rng.random() < 0.7:
            sh = [2 ** rng.randint(1, 9) for _ in range(D)]
        else:
            sh = [rng.choice([2, 3, 4, 5, 6, 8, 12, 16, 31, 32, 33, 64, 100, 128]) for _ in range(D)]
        n = int(np.prod(sh))
        if even and any(x % 2 for x in sh):
            continue
        if 8 <= n <= maxn:
            return sh


def _case(rng):
    kind = rng.choice(["ising", "ising", "potts", "z2", "cluster"])
    D = rng.randint(2, 4) if kind == "z2" else rng.randint(1, 4)
    c = {"kind": kind, "D": D, "kernel": rng.choice([O.KERNEL_METROPOLIS, O.KERNEL_HEAT_BATH]),
         "profile": rng.choice([O.PROFILE_F64, O.PROFILE_F32]), "seed": rng.randint(0, 10**9), "env": {}}
    if kind in ("ising", "potts"):
        c["shape"] = _shape(rng, D, 300000, True)
        c["q"] = 2 if kind == "ising" else rng.randint(3, 6)
        c["beta"] = rng.choice([0.05, 0.2216546, 0.3, 0.4406868, 0.7, 1.5, 4.0])
        c["j"] = rng.choice([1.0, 1.0, -0.7, 0.3])
        c["h"] = rng.choice([0.0, 0.0, 0.1, -0.25, 0.6]) if kind == "ising" else 0.0
        c["offsets"] = [rng.choice([0.0, 0.0, 0.2, -0.3]) for _ in range(c["q"])] if kind == "potts" else None
        c["env"] = {"PLQ_X16U": rng.choice([None, "0"]), "PLQ_SMALL_CHAIN": rng.choice([None, None, "0"]),
                    "PLQ_PASS_BLOCKS": rng.choice([None, None, "8", "64", "512"])}
        c["cold"] = rng.random() < 0.2
    elif kind == "z2":
        c["shape"] = _shape(rng, D, 40000, True)
        c["beta"] = rng.choice([0.1, 0.44, 0.7, 1.2])
    else:
        c["shape"] = _shape(rng, D, 200000, False)
        c["q"] = rng.randint(2, 5)
        c["seeded"] = rng.random() < 0.5
        c["forced"] = (rng.random() < 0.5) if (c["seeded"] or c["q"] > 2) else False
        c["gap"] = rng.choice([1.0, 2.0])
        c["beta"] = rng.choice([0.1, 0.4, 0.7, 1.0050525, 1.5])
        c["env"] = {"PLQ_SMALL_CLUSTER": rng.choice([None, "0"]), "PLQ_WOLFF_GROW": rng.choice([None, "0", "1", "2"]),
                    "PLQ_WOLFF_LEVELS": rng.choice([None, "3", "40"])}
    return c


CASES = [_case(random.Random(1000 + k)) for k in range(300)]


@pytest.mark.parametrize("block", range(6))
def test_random_configurations_are_bit_exact(gpu, monkeypatch, block):
    for c in CASES[block * 50:(block + 1) * 50]:
        for k in SWITCHES:
            monkeypatch.delenv(k, raising=False)
        for k, v in c["env"].items():
            if v is not None:
                monkeypatch.setenv(k, v)
        shape, seed = c["shape"], c["seed"]
        n = int(np.prod(shape))
        if c["kind"] in ("ising", "potts"):
            q = c["q"]
            start = np.zeros(n, np.uint8) if c["cold"] else np.random.default_rng(seed).integers(0, q, n).astype(np.uint8)
            if c["kind"] == "ising":
                ch = gpu.ising_gpu_chain(shape, c["j"], c["h"], c["beta"], seed, start, kernel=c["kernel"],
                                         profile=c["profile"])
                d = O.local_desc(shape, kernel=c["kernel"], profile=c["profile"], beta=c["beta"], j=c["j"], h=c["h"],
                                 seed=seed)
            else:
                ch = gpu.potts_gpu_chain(shape, q, c["j"], c["offsets"], c["beta"], seed, start, kernel=c["kernel"],
                                         profile=c["profile"])
                d = O.local_desc(shape, model=O.MODEL_POTTS, q=q, kernel=c["kernel"], profile=c["profile"],
                                 beta=c["beta"], j=c["j"], offsets=c["offsets"], seed=seed)
        elif c["kind"] == "z2":
            start = np.random.default_rng(seed).integers(0, 2, n * len(shape)).astype(np.uint8)
            ch = gpu.gauge_gpu_chain(shape, 1.0, c["beta"], seed, start, kernel=c["kernel"], profile=c["profile"])
            d = O.local_desc(shape, model=O.MODEL_Z2, kernel=c["kernel"], profile=c["profile"], beta=c["beta"],
                             seed=seed)
        else:
            q = c["q"]
            start = np.random.default_rng(seed).integers(0, q, n).astype(np.uint8)
            ext = gpu.Extent.SEEDED if c["seeded"] else gpu.Extent.ALL
            rel = gpu.Relabel.FORCED_CHANGE if c["forced"] else gpu.Relabel.REDRAW
            ch = gpu.GpuClusterChain.new(shape, q, c["gap"], ext, rel, c["beta"], seed, start, profile=c["profile"])
            d = O.cluster_desc(shape, q=q, seeded=c["seeded"], forced=c["forced"], profile=c["profile"],
                               beta=c["beta"], gap=c["gap"], seed=seed)
        st = start
        for t in range(3):
            ch.advance(1)
            if c["kind"] == "cluster":
                st, roots = O.cluster_sweeps(d, st, t, 1)
            else:
                st = O.local_sweeps(d, st, t, 1)
            got = ch.read()
            assert np.array_equal(got, st), (c, t, int((got != st).sum()))
        if c["kind"] == "cluster":
            assert np.array_equal(ch.roots(), roots), c
        ch.close()
