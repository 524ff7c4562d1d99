"""Cholesky of the Schur matrix through the C-ABI on matrices whose factorisation is known exactly.

B = L D L^T with L unit lower triangular (dyadic entries, |l_ij| <= 1/(4m), so kappa(L) <= 5/3) and
D diagonal: the k-th pivot of B's Cholesky is d_k and the k-th leading principal minor is
d_1 ... d_k.  With d_k = -1 at one position f and d_i in [1, 3] elsewhere, the first non-positive
pivot is f, which is the info that Rpotrf2 reports (mplapack/Rpotrf2.cpp:86-97) and that
choleskyFactorWithAdjust turns into FAILURE (sdpa_linear.cpp:622-630).  The tests check:
  - sdpg_potrf returns info = f for failures in the first block, at and after a block boundary,
    in a later block and at the last pivot, and the oracle's Rpotrf reports the same f on the
    same matrix;
  - the factor columns finished before the failing diagonal block are the factor of B there
    (L~ L~^T = B on those columns, exact integer arithmetic);
  - a positive definite B on the same context afterwards factors with info = 0 and
    L~ L~^T = B (a failure leaves no state behind: the pivot flag is reset).
"""
import ctypes
import math
from fractions import Fraction

import numpy as np
import pytest

import exact
import oracle_lib
import sdpg


def ldl(m, fail_at, seed):
    """exact B = L D L^T (Fractions, column-major flattened) with d_{fail_at} = -1 (1-based; 0: none)"""
    rng = np.random.default_rng(seed)
    q = 1 << 14
    lim = max(1, q // (4 * m))
    Lm = np.eye(m, dtype=np.int64) * q
    for i in range(m):
        Lm[i, :i] = rng.integers(-lim, lim + 1, size=i)
    d = rng.integers(1, 4, size=m).astype(np.int64)
    if fail_at:
        d[fail_at - 1] = -1
    Lo = Lm.astype(object)
    B = (Lo * d.astype(object)).dot(Lo.T)  # scaled by q^2
    return [Fraction(int(B[i, j]), q * q) for j in range(m) for i in range(m)]


def oracle_info(Bp, m, L):
    lib = oracle_lib.lib()
    u32p = ctypes.POINTER(ctypes.c_uint32)
    lib.orc_potrf.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, u32p, u32p, ctypes.POINTER(ctypes.c_int)]
    a = np.ascontiguousarray(Bp, dtype=np.uint32)
    info = ctypes.c_int()
    assert lib.orc_potrf(256, m, L, a.ctypes.data_as(u32p), None, ctypes.byref(info)) == 0
    return info.value


def factor_residual(Lp, Bp, m, L, ncols):
    """max over the first ncols columns of |(L L^T - B)_ij| / sqrt(B_ii B_jj), exactly"""
    Li, EL = exact.to_ints(Lp, L)
    Bi, EB = exact.to_ints(Bp, L)
    Lm = Li.reshape(m, m, order="F")
    Lm = np.where(np.tril(np.ones((m, m), dtype=bool)), Lm, 0).astype(object)
    Bm = Bi.reshape(m, m, order="F")
    R = Lm.dot(Lm[:ncols, :].T)  # m x ncols, exponent 2 EL
    a, b, E = exact.align(R.reshape(-1), 2 * EL, Bm[:, :ncols].reshape(-1), EB)
    d = (a - b).reshape(m, ncols)
    bd = np.abs(exact.ldexp_obj(np.diag(Bm), EB))
    worst = 0.0
    for j in range(ncols):
        for i in range(m):
            if d[i, j]:
                worst = max(worst, abs(float(Fraction(int(d[i, j])) * Fraction(2) ** E)) / math.sqrt(bd[i] * bd[j]))
    return worst


@pytest.mark.gpu
@pytest.mark.parametrize("gemm", [sdpg.GEMM_PLAIN, sdpg.GEMM_OZAKI], ids=["plain", "ozaki"])
def test_potrf_reports_first_failing_pivot(gemm, workload):
    ctx = sdpg.Context(workload(1), 256, gemm=gemm)  # m = 351
    try:
        m, L = ctx.m, ctx.L
        P = 32 * L if gemm == sdpg.GEMM_PLAIN else 8 * (4 * L - 4) - 4
        for fail_at in (1, 5, 32, 33, 81, 200, m):
            print(f"case gemm={gemm} m={m} fail_at={fail_at}", flush=True)
            Bp = sdpg.pack(ldl(m, fail_at, seed=fail_at), L)
            ctx.put(sdpg.SDPG_B, Bp)
            info = ctx.potrf()
            assert info == fail_at, (info, fail_at)
            assert oracle_info(Bp, m, L) == fail_at
            done = ((fail_at - 1) // 32) * 32  # columns finished before the failing diagonal block
            if done:
                r = factor_residual(ctx.get(sdpg.SDPG_L), Bp, m, L, done)
                print(f"  first {done} columns: scaled residual {r:.3e}", flush=True)
                assert r <= 2.0 ** -(P - 16) * m, r
        print(f"case gemm={gemm} m={m} positive definite after the failures", flush=True)
        Bp = sdpg.pack(ldl(m, 0, seed=99), L)
        ctx.put(sdpg.SDPG_B, Bp)
        assert ctx.potrf() == 0
        assert oracle_info(Bp, m, L) == 0
        r = factor_residual(ctx.get(sdpg.SDPG_L), Bp, m, L, m)
        print(f"  all columns: scaled residual {r:.3e}", flush=True)
        assert r <= 2.0 ** -(P - 16) * m, r
    finally:
        ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec,rows", [(256, 12), (256, 7), (448, 5), (512, 0)],
                         ids=["256-R12", "256-R7", "448-R5", "512"])
def test_potrf_wide_workgroups_and_precisions(prec, rows, workload, monkeypatch):
    """the block-column kernel with more panel rows per workgroup than the max-cut sizes use (R = 5 .. 12:
    up to 880 items, 1024 threads, 44 rows for the row-scale lanes) and at 448 / 512 bits (16 / 18 limbs of
    32 bits, 18 / 20 of 29 bits in the elimination): info of a failing pivot, the factor before it, and a
    positive definite factor, exactly"""
    if rows:
        monkeypatch.setenv("SDPG_COL_R", str(rows))
    ctx = sdpg.Context(workload(1), prec, gemm=sdpg.GEMM_PLAIN)  # m = 351
    try:
        m, L = ctx.m, ctx.L
        P = 32 * L
        fail_at = 100
        print(f"case prec={prec} R={rows} m={m} fail_at={fail_at}", flush=True)
        Bp = sdpg.pack(ldl(m, fail_at, seed=7), L)
        ctx.put(sdpg.SDPG_B, Bp)
        assert ctx.potrf() == fail_at
        r = factor_residual(ctx.get(sdpg.SDPG_L), Bp, m, L, 96)
        print(f"  first 96 columns: scaled residual {r:.3e}", flush=True)
        assert r <= 2.0 ** -(P - 16) * m, r
        print(f"case prec={prec} R={rows} m={m} positive definite", flush=True)
        Bp = sdpg.pack(ldl(m, 0, seed=8), L)
        ctx.put(sdpg.SDPG_B, Bp)
        assert ctx.potrf() == 0
        r = factor_residual(ctx.get(sdpg.SDPG_L), Bp, m, L, m)
        print(f"  all columns: scaled residual {r:.3e}", flush=True)
        assert r <= 2.0 ** -(P - 16) * m, r
    finally:
        ctx.close()
