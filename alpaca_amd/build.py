"""Build the gfx950 HIP library and the CPU checker libraries in-tree.

``build_hip()`` compiles ``alpaca_amd/csrc/*.hip`` into
``alpaca_amd/libalpaca_hip.so`` with hipcc for gfx950 only.
``build_oracle()`` runs ``make -C oracle`` (the C restatement used by the
tests) and, where the reference source tree is present, ``make -C oracle ref``
(the harness linking the reference's own translation units).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "alpaca_amd")
CSRC = os.path.join(PKG, "csrc")
HIP_LIB = os.path.join(PKG, "libalpaca_hip.so")
REFERENCE = "/root/reference"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off keeps a*b+c unfused so the FP64 path reproduces the
# reference's x86 build (no FMA) bit for bit.  Packed FP32 instructions are
# disabled for every HIP translation unit.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

HIP_SOURCES = ["alpaca_kernels.hip", "ensemble.hip", "multiresolution.hip", "lts.hip"]

def _run(cmd, cwd=None, verbose=False):
    if verbose:
        print("+", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return res


def build_hip(verbose=False):
    """Compile every HIP translation unit (in parallel) and link the library."""
    from concurrent.futures import ThreadPoolExecutor

    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        jobs.append(([HIPCC, *HIP_FLAGS, "-c", "-o", obj, os.path.join(CSRC, src)], obj))
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        list(ex.map(lambda j: _run(j[0], verbose=verbose), jobs))
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *[o for _, o in jobs]],
         verbose=verbose)
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_host_selftest(verbose=False):
    """Command-line driver of the C++ host adapters (host code only)."""
    src = os.path.join(CSRC, "host", "selftest.cpp")
    out = os.path.join(PKG, "alpaca_selftest")
    _run([HIPCC, "-O2", "-std=c++17", "-Wno-unused-command-line-argument", "-o", out + ".tmp",
          src, "-L" + PKG, "-lalpaca_hip", "-Wl,-rpath,$ORIGIN"], verbose=verbose)
    os.replace(out + ".tmp", out)
    return out


def build_oracle(verbose=False):
    make = shutil.which("make") or "make"
    _run([make, "-C", os.path.join(ROOT, "oracle"), "-j8"], verbose=verbose)
    if os.path.isdir(os.path.join(REFERENCE, "src")):
        _run([make, "-C", os.path.join(ROOT, "oracle"), "-j8", "ref", f"REF={REFERENCE}"],
             verbose=verbose)


def build_all(verbose=False):
    build_hip(verbose)
    build_host_selftest(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
