"""sdpg_rgemm_noalloc: the Rgemm entry point that allocates nothing (its device memory is a static
array of the library), checked against sdpg_rgemm, which allocates per call.

The two entry points run the same kernels on the same inputs, so their outputs must be identical bit
for bit; both are also held to the exact-product bars of the GEMM tests.  Each form is called twice
in a row through the no-allocation entry point, the second time with smaller operands (the array
keeps the first call's contents, which the second call must not depend on), and a request whose
operands exceed the array is rejected with SDPG_EARG, leaving C untouched.
"""
import random
from fractions import Fraction

import numpy as np
import pytest

import sdpg
from test_gpu_gemm import FORMS, FORM_IDS, L, PREC, _check, _exact, _pack2, _rand_matrix

SDPG_EARG = -1


@pytest.mark.gpu
@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
def test_noalloc_matches_allocating_entry_point(form):
    mode, slices, bits_lost = form
    shapes = [(120, 90, 150, False, True, 1), (37, 21, 24, True, False, 0)]  # larger, then smaller
    for m, n, k, ta, tb, op in shapes:
        print(f"case noalloc mode={mode} slices={slices} shape={(m, n, k, ta, tb, op)}", flush=True)
        rnd = random.Random(hash((m, n, k, mode, slices)) & 0xFFFF)
        Av = _rand_matrix(rnd, k if ta else m, m if ta else k)
        Bv = _rand_matrix(rnd, n if tb else k, k if tb else n)
        Cv = _rand_matrix(rnd, m, n)
        Ap, Bp, Cp = _pack2(Av), _pack2(Bv), _pack2(Cv)
        ref, _ = sdpg.rgemm(PREC, Ap, Bp, Cp, ta, tb, op, mode, slices)
        got, ms = sdpg.rgemm(PREC, Ap, Bp, Cp, ta, tb, op, mode, slices, noalloc=True)
        assert ms >= 0
        assert np.array_equal(got, ref), "sdpg_rgemm_noalloc differs from sdpg_rgemm"
        A, B, ex = _exact(Av, Bv, Cv, ta, tb, op)
        _check(got, A, B, Cv, ex, op, bits_lost)


@pytest.mark.gpu
def test_noalloc_lower_update_twice():
    """the Cholesky trailing-update shape, C -= P P^T on the lower triangle, twice in a row"""
    for m, k in ((90, 40), (33, 8)):
        rnd = random.Random(m)
        Pv = _rand_matrix(rnd, m, k, espread=4)
        Cv = _rand_matrix(rnd, m, m, espread=4)
        Pp, Cp = _pack2(Pv), _pack2(Cv)
        for mode in (sdpg.GEMM_PLAIN, sdpg.GEMM_OZAKI):
            print(f"case noalloc-lower mode={mode} m={m} k={k}", flush=True)
            ref, _ = sdpg.rgemm(PREC, Pp, Pp, Cp, False, True, 1, mode, 0, lower=True)
            got, _ = sdpg.rgemm(PREC, Pp, Pp, Cp, False, True, 1, mode, 0, lower=True, noalloc=True)
            assert np.array_equal(got, ref)


@pytest.mark.gpu
def test_noalloc_rejects_oversize():
    """operands of 1500 x 1500 at 256 bits need 3 x 99 MB on the device: beyond the 128 MiB array"""
    n = 1500
    Z = np.zeros((n, n, L + 2), dtype=np.uint32)
    Z[:, :, 0] = np.uint32(sdpg.EZERO & 0xFFFFFFFF)
    C = Z.copy()
    C[0, 0] = sdpg.frac_to_words(Fraction(3), L)
    print("case noalloc-oversize", flush=True)
    with pytest.raises(sdpg.SdpgError) as ei:
        sdpg.rgemm(PREC, Z, Z, C, gemm=sdpg.GEMM_PLAIN, noalloc=True)
    assert ei.value.rc == SDPG_EARG
    assert "scratch" in str(ei.value)
    # a normal call still works afterwards
    A = _pack2([[Fraction(2)]])
    got, _ = sdpg.rgemm(PREC, A, A, None, noalloc=True)
    assert sdpg.words_to_frac(got[0, 0], L) == 4
