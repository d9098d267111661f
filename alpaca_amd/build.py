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
    "-shared",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

HIP_SOURCES = ["alpaca_kernels.hip"]


def _run(cmd, cwd=None, verbose=False):
    if verbose:
        print("+", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return res


def build_hip(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, *srcs], verbose=verbose)
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
