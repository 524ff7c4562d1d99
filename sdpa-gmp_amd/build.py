"""Build the native parts of the package in-tree.

Outputs (git-ignored, but shipped with the tree):
  sdpa-gmp_amd/lib/libsdpg.so     device engine (HIP, gfx950) + C-ABI + host driver
  sdpa-gmp_amd/lib/sdpa_amd       command-line solver (sdpa_gmp-compatible flags)
  sdpa-gmp_amd/lib/libmpxprobe.so host-only arithmetic probe used by CPU tests
Environment: SDPG_FAST_BUILD=1 compiles only the 256-bit (10-limb) instantiation; MAX_JOBS bounds the
parallel compiles.
"""
from __future__ import annotations

import concurrent.futures
import os
import re
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
# hipcc flags for all device code: no packed FP32 instructions (some of their forms have given wrong
# results on this target without any error)
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


def limb_counts(fast: bool) -> list[int]:
    """The limb counts of SDPG_FOR_EACH_L in limbs.h (the fast build keeps only L = 10)."""
    with open(os.path.join(CSRC, "limbs.h")) as f:
        text = f.read()
    body = text.split("#ifdef SDPG_FAST_BUILD", 1)[1]
    fast_part, full_part = body.split("#else", 1)
    part = fast_part if fast else full_part.split("#endif", 1)[0]
    return [int(x) for x in re.findall(r"X\((\d+)\)", part)]


def _run_to(cmd):
    """Run a compile whose last two arguments are '-o <obj>', writing <obj>.tmp first."""
    out = cmd[-1]
    dt = _run([*cmd[:-1], out + ".tmp"])
    os.replace(out + ".tmp", out)
    return dt


def build(fast: bool | None = None, verbose: bool = True) -> None:
    if fast is None:
        fast = os.environ.get("SDPG_FAST_BUILD", "0") == "1"
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    defs = ["-DSDPG_FAST_BUILD"] if fast else []
    inc = ["-I", CSRC, "-I", INC]
    # one translation unit per limb count (engine_inst.hip), the C-ABI (engine.hip) and the host
    # driver compile in parallel; objects are written to temporary names and renamed when whole
    jobs = []
    objs = []
    eng_o = os.path.join(OBJ, "engine.o")
    jobs.append(("engine.hip", [HIPCC, *HIP_FLAGS, *defs, *inc, "-c", os.path.join(CSRC, "engine.hip"), "-o", eng_o]))
    objs.append(eng_o)
    for L in limb_counts(fast):
        o = os.path.join(OBJ, f"engine_L{L}.o")
        jobs.append((f"engine_inst.hip (L={L})", [HIPCC, *HIP_FLAGS, *defs, *inc, f"-DSDPG_INST_L={L}", "-c",
                                                  os.path.join(CSRC, "engine_inst.hip"), "-o", o]))
        objs.append(o)
    hs_o = os.path.join(OBJ, "host_solver.o")
    jobs.append(("host_solver.cpp", ["g++", *CXX_FLAGS, *defs, *inc, "-c", os.path.join(CSRC, "host_solver.cpp"),
                                     "-o", hs_o]))
    objs.append(hs_o)
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    t0 = time.time()
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        futs = {pool.submit(_run_to, cmd): name for name, cmd in jobs}
        for f in concurrent.futures.as_completed(futs):
            dt = f.result()
            if verbose:
                print(f"[build] {futs[f]}: {dt:.1f}s", flush=True)
    so = os.path.join(LIB, "libsdpg.so")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-fopenmp", *objs, "-o", so + ".tmp"])
    os.replace(so + ".tmp", so)
    if verbose:
        print(f"[build] libsdpg.so linked, {time.time() - t0:.1f}s wall", flush=True)
    cli = os.path.join(LIB, "sdpa_amd")
    _run(["g++", *CXX_FLAGS, *inc, os.path.join(CSRC, "sdpa_amd_main.cpp"), "-o", cli, "-L", LIB, "-lsdpg",
          "-Wl,-rpath,$ORIGIN"])
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", *inc, os.path.join(CSRC, "mpx_probe.cpp"), "-o",
          os.path.join(LIB, "libmpxprobe.so")])
    if verbose:
        print("[build] done:", so, flush=True)


if __name__ == "__main__":
    build(fast=("--fast" in sys.argv) or None)
