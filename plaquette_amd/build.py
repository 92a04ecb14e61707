"""Build ``plaquette_amd/libplaquette_hip.so`` for gfx950 with hipcc.

Every source is compiled for ``--offload-arch=gfx950`` only, with
``-ffp-contract=off`` (the host-side threshold arithmetic must round each
operation once, as the reference's Rust and WGSL do) and with packed FP32
instructions disabled in device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libplaquette_hip.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = [
    "plq_abi.hip",
    "plq_site.hip",
    "plq_link.hip",
    "plq_cluster.hip",
    "plq_cluster_small.hip",
    "plq_ensemble.hip",
    "plq_ensemble_abi.hip",
    "plq_small.hip",
    "plq_pmode.hip",
    "plq_block.hip",
    "plq_block_abi.hip",
    "plq_measure.hip",
    "plq_measure_abi.hip",
    "plq_wolff.hip",
    "plq_thresholds.cpp",
]

ARCH = os.environ.get("PLQ_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the HIP extension cannot be built")
    return path


def common_flags() -> list[str]:
    return [
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-ffp-contract=off",
        "-fno-fast-math",
        "-Wall",
        "-Wno-unused-function",
        "-I" + os.path.join(REPO, "include"),
        "-I" + CSRC,
    ]


def compile_one(src: str) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [hipcc(), "-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
               "-packed-fp32-ops", "-c", path, "-o", obj] + common_flags()
    else:
        cmd = [hipcc(), "-x", "c++", "-c", path, "-o", obj] + common_flags()
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"compiling {src} failed")
    # hipcc warns on the host pass about the device-only target feature; keep
    # the rest visible
    noise = "'-packed-fp32-ops' is not a recognized feature"
    for line in (res.stdout + res.stderr).splitlines():
        if noise not in line and "warning generated" not in line and line.strip():
            sys.stderr.write(line + "\n")
    return obj


def build(verbose: bool = False) -> str:
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), 8)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("linking libplaquette_hip.so failed")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
