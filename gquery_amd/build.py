"""Compile the native library in-tree: gquery_amd/libgquery_amd.so (CPU builder + HIP kernels for
gfx950 + C ABI).  Run `python -m gquery_amd.build` or call build()."""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libgquery_amd.so"
OBJ = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GQ_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
# Translation units and their floating-point contraction mode.  The plain kernels reproduce the
# CPU restatement of the reference traversal bit-for-bit, so they must not contract a*b+c.
HIP_UNITS = {
    "plain.hip": ["-ffp-contract=off"],
    "baseline.hip": ["-ffp-contract=off"],
    "sample.hip": ["-ffp-contract=off"],
    "engine.hip": ["-ffp-contract=off"],
    "gq_api.hip": [],
    "silprep.hip": ["-ffp-contract=off"],
}
CXX_UNITS = {
    # The builder must evaluate float expressions exactly as written (no FMA contraction).
    "builder.cpp": ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math"],
    "wide.cpp": ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
                 "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"],
}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(map(str, cmd)) + "\n" + r.stdout)
    return r.stdout


def build(verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    jobs = []
    for src, extra in HIP_UNITS.items():
        if not (CSRC / src).exists():
            continue
        o = OBJ / (src + ".o")
        jobs.append([HIPCC, "-c", str(CSRC / src), "-o", str(o)] + HIP_FLAGS + extra)
    for src, extra in CXX_UNITS.items():
        if not (CSRC / src).exists():
            continue
        o = OBJ / (src + ".o")
        jobs.append(["g++", "-c", str(CSRC / src), "-o", str(o), "-Wall"] + extra)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    objs = [str(OBJ / (s + ".o")) for s in list(HIP_UNITS) + list(CXX_UNITS) if (CSRC / s).exists()]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp)] + objs)
    os.replace(tmp, LIB)
    # the VALU issue-rate probe behind bench.py's roofline peak (a separate program)
    probe = CSRC / "valu_probe.hip"
    if probe.exists():
        _run([HIPCC, str(probe), "-o", str(PKG / "valu_probe")] + [f for f in HIP_FLAGS if f != "-fPIC"])
    return LIB


def clean():
    shutil.rmtree(OBJ, ignore_errors=True)
    if LIB.exists():
        LIB.unlink()


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
