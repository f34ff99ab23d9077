"""Build the native parts in-tree.

* ``ablsteadyrans_amd/libgust_amd.so``: the HIP library (every ``csrc/*.hip``,
  plus the host-only ``csrc/*.cpp``)
  for gfx950, C-ABI declared in ``include/gust_amd.h``.
* ``oracle/liboracle.so``: the FP64 C restatement of the reference (test
  infrastructure; built here so the checker exists on the GPU host too).

Flags: no floating-point contraction anywhere (the reference's CPU build
rounds every operation separately; with contraction off the FP64 GPU results
equal it bit for bit), and packed FP32 instructions disabled on the device.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ablsteadyrans_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libgust_amd.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))


def build_hip(verbose: bool = False) -> str:
    """Compile every source in parallel (one hipcc per file; the box and this
    container both have several cores), then link."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(os.path.join(PKG, "_build"), exist_ok=True)
    cmds, objs = [], []
    for s in sorted(glob.glob(os.path.join(CSRC, "*.hip"))):
        o = os.path.join(PKG, "_build", os.path.basename(s) + ".o")
        cmds.append([HIPCC] + [f for f in HIP_FLAGS if f != "-shared"] + [
            "-c", "-I", os.path.join(ROOT, "include"), "-I", CSRC, s, "-o", o])
        objs.append(o)
    # Host-only C++ (the steady criterion): the host compiler, same rounding.
    for s in sorted(glob.glob(os.path.join(CSRC, "*.cpp"))):
        o = os.path.join(PKG, "_build", os.path.basename(s) + ".o")
        cmds.append(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-c",
                     "-I", os.path.join(ROOT, "include"), "-I", CSRC, s, "-o", o])
        objs.append(o)
    if verbose:
        for c in cmds:
            print(" ".join(c))
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> str:
    srcs = sorted(glob.glob(os.path.join(ORACLE_DIR, "*.c")))
    tmp = ORACLE_LIB + ".tmp"
    # x86-64-v3 (AVX2): the build host and the GPU hosts both have it; the
    # separate roundings (-ffp-contract=off) keep every result IEEE's
    cmd = ["gcc", "-O3", "-std=c11", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
           "-shared", "-Wall", "-Wno-unused-function", *srcs, "-o", tmp, "-lm"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build_all(verbose: bool = False) -> None:
    build_oracle(verbose)
    build_hip(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
