"""Build the native parts of the package in-tree.

Outputs (git-ignored, but shipped with the tree):
  sdpa-gmp_amd/lib/libsdpg.so     device engine (HIP, gfx950) + C-ABI + host driver
  sdpa-gmp_amd/lib/sdpa_amd       command-line solver (sdpa_gmp-compatible flags)
  sdpa-gmp_amd/lib/libmpxprobe.so host-only arithmetic probe used by CPU tests
Environment: SDPG_FAST_BUILD=1 compiles only the 256-bit (8-limb) instantiation.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SDPG_ARCH", "gfx950")
# hipcc flags for all device code: no packed FP32 instructions (see DESIGN notes on the target)
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-pass-failed"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp"]


def _run(cmd, cwd=None):
    t = time.time()
    p = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return time.time() - t


def build(fast: bool | None = None, verbose: bool = True) -> None:
    if fast is None:
        fast = os.environ.get("SDPG_FAST_BUILD", "0") == "1"
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    defs = ["-DSDPG_FAST_BUILD"] if fast else []
    inc = ["-I", CSRC, "-I", INC]
    steps = []
    eng_o = os.path.join(OBJ, "engine.o")
    steps.append(("engine.hip", [HIPCC, *HIP_FLAGS, *defs, *inc, "-c", os.path.join(CSRC, "engine.hip"), "-o", eng_o]))
    hs_o = os.path.join(OBJ, "host_solver.o")
    steps.append(("host_solver.cpp", ["g++", *CXX_FLAGS, *defs, *inc, "-c", os.path.join(CSRC, "host_solver.cpp"), "-o", hs_o]))
    so = os.path.join(LIB, "libsdpg.so")
    steps.append(("libsdpg.so", [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-fopenmp", eng_o, hs_o, "-o", so + ".tmp"]))
    for name, cmd in steps:
        dt = _run(cmd)
        if verbose:
            print(f"[build] {name}: {dt:.1f}s", flush=True)
    os.replace(so + ".tmp", so)
    cli = os.path.join(LIB, "sdpa_amd")
    _run(["g++", *CXX_FLAGS, *inc, os.path.join(CSRC, "sdpa_amd_main.cpp"), "-o", cli, "-L", LIB, "-lsdpg",
          "-Wl,-rpath,$ORIGIN"])
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", *inc, os.path.join(CSRC, "mpx_probe.cpp"), "-o",
          os.path.join(LIB, "libmpxprobe.so")])
    if verbose:
        print("[build] done:", so, flush=True)


if __name__ == "__main__":
    build(fast=("--fast" in sys.argv) or None)
