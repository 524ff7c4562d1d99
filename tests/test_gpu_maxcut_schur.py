"""The Schur matrix of the max-cut workload (n = m = 1000, A_i = e_i e_i^T) against its closed form.

For A_i = e_i e_i^T every pair takes formula F3 (sdpa_newton.cpp:939-1005) and
B_ij = Tr(e_i e_i^T X e_j e_j^T inv(Z)) = X_ij inv(Z)_ij: the Hadamard product X o inv(Z), one
rounded product per entry.  The test puts a random symmetric X and inv(Z) with mantissas of full
width and exponents spread over 80 binades on the device, runs sdpg_schur, and compares every
diagonal entry and 30000 random entries of B with the exact product: within one unit in the last
limb (2^-(32L-1) relative).
"""
from fractions import Fraction

import numpy as np
import pytest

import sdpg


def random_symmetric(n, L, rng):
    """packed n x n symmetric matrix (column-major) of random full-width values"""
    iu, ju = np.triu_indices(n)
    cnt = len(iu)
    w = np.zeros((cnt, L + 2), dtype=np.uint32)
    w[:, 0] = rng.integers(-40, 41, size=cnt).astype(np.int32).view(np.uint32)
    w[:, 1] = rng.integers(0, 2, size=cnt)
    w[:, 2:] = rng.integers(0, 1 << 32, size=(cnt, L), dtype=np.uint64).astype(np.uint32)
    w[:, L + 1] |= np.uint32(1 << 31)
    out = np.zeros((n * n, L + 2), dtype=np.uint32)
    out[iu + n * ju] = w
    out[ju + n * iu] = w
    return out


@pytest.mark.gpu
def test_maxcut_schur_is_hadamard_product(workload):
    ctx = sdpg.Context(workload(2), 256, gemm=sdpg.GEMM_OZAKI)
    try:
        n, L = ctx.m, ctx.L
        assert ctx.nd == n * n
        rng = np.random.default_rng(2024)
        X = random_symmetric(n, L, rng)
        iZ = random_symmetric(n, L, rng)
        y = np.zeros((n, L + 2), dtype=np.uint32)
        y[:, 0] = np.uint32(sdpg.EZERO & 0xFFFFFFFF)
        print(f"case max-cut n={n} L={L}: sdpg_schur on a random point", flush=True)
        ctx.set_point(X, X, iZ, y)
        ctx.schur()
        B = ctx.get(sdpg.SDPG_B)
    finally:
        ctx.close()
    pick = np.unique(np.concatenate([np.arange(n) * (n + 1), rng.integers(0, n * n, size=30000)]))
    ulp = Fraction(1, 2 ** (32 * L - 1))
    worst = Fraction(0)
    for p in pick:
        got = sdpg.words_to_frac(B[p], L)
        want = sdpg.words_to_frac(X[p], L) * sdpg.words_to_frac(iZ[p], L)
        err = abs(got - want) / abs(want)
        worst = max(worst, err)
    print(f"  {len(pick)} entries, worst relative error {float(worst):.3e} (2^-(32L-1) = {float(ulp):.3e})")
    assert worst <= ulp
