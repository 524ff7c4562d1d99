"""ctypes access to the CPU oracle (oracle/_build/liborc.so) for tests and fixture scripts.

The oracle is test infrastructure: it is the checker, never the thing measured.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_DIR = os.path.join(ROOT, "oracle")
ORC_LIB = os.path.join(ORC_DIR, "_build", "liborc.so")
ORC_CLI = os.path.join(ORC_DIR, "_build", "orc_cli")


class OrcResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("phase", ctypes.c_int), ("phase_name", ctypes.c_char * 16),
                ("mu", ctypes.c_double), ("relgap", ctypes.c_double), ("gap", ctypes.c_double),
                ("objP", ctypes.c_char * 80), ("objD", ctypes.c_char * 80), ("relgap_s", ctypes.c_char * 80),
                ("pfeas_s", ctypes.c_char * 80), ("dfeas_s", ctypes.c_char * 80), ("t_bmat", ctypes.c_double),
                ("t_chol", ctypes.c_double), ("t_total", ctypes.c_double), ("objP70", ctypes.c_char * 112),
                ("objD70", ctypes.c_char * 112)]


class OrcParam(ctypes.Structure):
    _fields_ = [("maxIteration", ctypes.c_int), ("epsilonStar", ctypes.c_double), ("lambdaStar", ctypes.c_double),
                ("omegaStar", ctypes.c_double), ("lowerBound", ctypes.c_double), ("upperBound", ctypes.c_double),
                ("betaStar", ctypes.c_double), ("betaBar", ctypes.c_double), ("gammaStar", ctypes.c_double),
                ("epsilonDash", ctypes.c_double), ("precision", ctypes.c_int)]


_lib = None


def build_oracle() -> None:
    import subprocess
    subprocess.run(["make", "-s", "-C", ORC_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORC_LIB):
            build_oracle()
        L = ctypes.CDLL(ORC_LIB)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        i32p = ctypes.POINTER(ctypes.c_int32)
        L.orc_arith.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int32, ctypes.c_int32, u32p,
                                ctypes.c_int32, ctypes.c_int32, u32p, i32p, i32p, u32p]
        L.orc_parse.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, i32p, i32p, u32p]
        L.orc_format.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int32, ctypes.c_int32, u32p, ctypes.c_int,
                                 ctypes.c_char_p, ctypes.c_int]
        L.orc_newton_step.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.orc_solve.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p,
                                ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.orc_param_default.argtypes = [ctypes.c_void_p]
        L.orc_param_read.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.orc_time_hotpath.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.orc_time_point.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double)]
        L.orc_time_f1_sample.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def params(path: str | None = None) -> OrcParam:
    p = OrcParam()
    lib().orc_param_default(ctypes.byref(p))
    if path:
        if lib().orc_param_read(ctypes.byref(p), path.encode()) != 0:
            raise RuntimeError(f"cannot read {path}")
    return p


def solve(data: str, param: str | None = None, prec: int = 0, log: str | None = None, dump_dir: str | None = None,
          dump_iters=(), dump_L: int = 12) -> OrcResult:
    p = params(param)
    r = OrcResult()
    its = (ctypes.c_int * max(1, len(dump_iters)))(*dump_iters)
    rc = lib().orc_solve(data.encode(), ctypes.byref(p), prec, log.encode() if log else None,
                         dump_dir.encode() if dump_dir else None, its, len(dump_iters), dump_L, ctypes.byref(r))
    if rc < 0:
        raise RuntimeError(f"orc_solve rc={rc}")
    return r


def newton_step(data: str, prec: int, in_dir: str, out_dir: str, out_L: int) -> int:
    return lib().orc_newton_step(data.encode(), prec, in_dir.encode(), out_dir.encode(), out_L)


def time_hotpath(data: str, param: str | None, prec: int, warm: int, reps: int, threads: int):
    p = params(param)
    a, b = ctypes.c_double(), ctypes.c_double()
    rc = lib().orc_time_hotpath(data.encode(), ctypes.byref(p), prec, warm, reps, threads, ctypes.byref(a),
                                ctypes.byref(b))
    if rc < 0:
        raise RuntimeError(f"orc_time_hotpath rc={rc}")
    return a.value, b.value


def time_point(data: str, prec: int, x_path: str, invz_path: str, L: int, chol_n: int, threads: int):
    """seconds of Make_bMat on a dumped iterate and of the Cholesky of B's leading chol_n block"""
    a, b = ctypes.c_double(), ctypes.c_double()
    rc = lib().orc_time_point(data.encode(), prec, x_path.encode(), invz_path.encode(), L, chol_n, threads,
                              ctypes.byref(a), ctypes.byref(b))
    if rc < 0:
        raise RuntimeError(f"orc_time_point rc={rc}")
    return a.value, b.value


def time_f1_sample(data: str, prec: int, x_path: str, invz_path: str, L: int, nsample: int, threads: int):
    """seconds per constraint's F1 precompute and per pair's inner product (all-F1 block 0)"""
    a, b = ctypes.c_double(), ctypes.c_double()
    rc = lib().orc_time_f1_sample(data.encode(), prec, x_path.encode(), invz_path.encode(), L, nsample, threads,
                                  ctypes.byref(a), ctypes.byref(b))
    if rc < 0:
        raise RuntimeError(f"orc_time_f1_sample rc={rc}")
    return a.value, b.value


def read_dump(path: str) -> np.ndarray:
    """Dump file -> packed array (n, L+2) in the sdpg exchange layout [e, sign(0/1), limbs]."""
    raw = np.fromfile(path, dtype=np.int32)
    n, L = int(raw[0]), int(raw[1])
    sg = raw[2:2 + n]
    ex = raw[2 + n:2 + 2 * n]
    lb = raw[2 + 2 * n:2 + 2 * n + n * L].view(np.uint32).reshape(n, L)
    out = np.zeros((n, L + 2), dtype=np.uint32)
    out[:, 0] = np.where(sg == 0, -(1 << 29), ex).astype(np.int32).view(np.uint32)
    out[:, 1] = (sg < 0).astype(np.uint32)
    out[:, 2:] = lb
    return out
