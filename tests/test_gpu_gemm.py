"""Multiple-precision GEMM forms on the GPU (sdpg_rgemm) against exact integer products.

The reference reaches these products through Rgemm_NN_omp / Rgemm_NT_omp / Rsyrk
(mplapack/Rgemm_NN_omp.cpp:36-74, mplapack/Rgemm_NT_omp.cpp:36-76, mplapack/Rsyrk.cpp:157-177).
Inputs are dyadic, so the exact product sum is an integer computation.

Bars (per element, K terms, a_i = max_k |A[i,k]|, b_j = max_k |B[k,j]|):
  plain   |C - exact| <= K a_i b_j 2^-(32L-10)  + rounding of the result
  element |C - exact| <= K a_i b_j 2^-(32L-10)  + rounding
  ozaki   |C - exact| <= K a_i b_j 2^-(8S-12)   + rounding,  S = slices (default 4L-4)
and exact equality for small-integer operands (every form is exact there).
"""
import random
from fractions import Fraction

import numpy as np
import pytest

import sdpg

PREC = 256
L = sdpg.limbs_for_bits(PREC)


def _rand_matrix(rnd, r, c, espread=20, zero_rows=()):
    vals = []
    for i in range(r):
        row = []
        for j in range(c):
            if i in zero_rows or rnd.random() < 0.03:
                row.append(Fraction(0))
                continue
            M = rnd.getrandbits(32 * L) | (1 << (32 * L - 1))
            e = rnd.randint(-espread, espread)
            v = Fraction(M, 1 << (32 * L)) * Fraction(2) ** e
            row.append(-v if rnd.random() < 0.5 else v)
        vals.append(row)
    return vals


def _pack2(vals):
    r, c = len(vals), len(vals[0])
    out = np.zeros((r, c, L + 2), dtype=np.uint32)
    for i in range(r):
        for j in range(c):
            out[i, j] = sdpg.frac_to_words(vals[i][j], L)
    return out


def _exact(Av, Bv, Cv, transa, transb, op):
    """exact op(A) op(B) combined with C through scaled Python integers"""
    A = [list(x) for x in zip(*Av)] if transa else Av
    B = [list(x) for x in zip(*Bv)] if transb else Bv
    m, k, n = len(A), len(B), len(B[0])
    den = 1
    for M_ in (A, B):
        for row in M_:
            for v in row:
                den = max(den, v.denominator)
    Ai = np.array([[int(v * den) for v in row] for row in A], dtype=object)
    Bi = np.array([[int(v * den) for v in row] for row in B], dtype=object)
    P = Ai.dot(Bi) if k else np.zeros((m, n), dtype=object)
    out = [[Fraction(int(P[i, j]), den * den) for j in range(n)] for i in range(m)]
    if op == 1:
        out = [[Cv[i][j] - out[i][j] for j in range(n)] for i in range(m)]
    elif op == 2:
        out = [[Cv[i][j] + out[i][j] for j in range(n)] for i in range(m)]
    return A, B, out


def _check(got, A, B, Cv, ex, op, bits_lost, lower=False):
    m, n = len(ex), len(ex[0])
    k = len(B)
    G = [[sdpg.words_to_frac(got[i, j], L) for j in range(n)] for i in range(m)]
    amax = [max((abs(v) for v in A[i]), default=0) for i in range(m)]
    bmax = [max((abs(B[t][j]) for t in range(k)), default=0) for j in range(n)]
    ulp = Fraction(1, 1 << (32 * L - 2))
    worst = 0.0
    for i in range(m):
        for j in range(n):
            if lower and i < j:
                assert G[i][j] == Cv[i][j], "upper triangle written"
                continue
            tol = max(k, 1) * amax[i] * bmax[j] * Fraction(1, 1 << bits_lost) + abs(ex[i][j]) * ulp
            if op:
                tol += abs(Cv[i][j]) * ulp
            err = abs(G[i][j] - ex[i][j])
            if tol:
                worst = max(worst, float(err / tol))
            assert err <= tol, (i, j, float(err), float(tol))
    return worst


FORMS = [(sdpg.GEMM_PLAIN, 0, 32 * L - 10), (sdpg.GEMM_ELEM, 0, 32 * L - 10),
         (sdpg.GEMM_OZAKI, 0, 8 * (4 * L - 4) - 12), (sdpg.GEMM_OZAKI, 4 * L, 8 * (4 * L) - 12)]
FORM_IDS = ["plain", "element", "ozaki", "ozaki-wide"]


@pytest.mark.gpu
@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
@pytest.mark.parametrize("shape", [(70, 67, 37, False, True, 0), (33, 129, 16, True, False, 1),
                                   (1, 5, 1, False, False, 2), (96, 96, 32, False, True, 1),
                                   (50, 40, 300, False, True, 1), (17, 20, 2100, True, False, 2),
                                   (40, 36, 150, True, True, 1), (65, 97, 64, False, True, 1),
                                   (130, 33, 49, True, False, 2)],
                         ids=["NT-set", "TN-sub", "tiny-add", "NT-sub", "NT-sub-K300", "TN-add-K2100", "TT-sub-K150",
                              "NT-sub-K64", "TN-add-K49"])
def test_rgemm_random(form, shape):
    mode, slices, bits_lost = form
    m, n, k, ta, tb, op = shape
    print(f"case rgemm mode={mode} slices={slices} shape={shape}", flush=True)
    rnd = random.Random(hash((m, n, k, ta, tb, op)) & 0xFFFF)
    Av = _rand_matrix(rnd, k if ta else m, m if ta else k, zero_rows=(3,))
    Bv = _rand_matrix(rnd, n if tb else k, k if tb else n)
    Cv = _rand_matrix(rnd, m, n)
    A, B, ex = _exact(Av, Bv, Cv, ta, tb, op)
    got, _ = sdpg.rgemm(PREC, _pack2(Av), _pack2(Bv), _pack2(Cv), ta, tb, op, mode, slices)
    worst = _check(got, A, B, Cv, ex, op, bits_lost)
    print(f"  worst error / bound = {worst:.3e}")


@pytest.mark.gpu
@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
@pytest.mark.parametrize("k", [32, 200], ids=["K32", "K200"])
def test_rgemm_lower_update(form, k):
    """the Cholesky trailing update shape (K = 32) and the Schur inner-product shape (long K):
    C -= P P^T on the lower triangle only"""
    mode, slices, bits_lost = form
    rnd = random.Random(7 + k)
    m = 75
    print(f"case rgemm-lower mode={mode} slices={slices} k={k}", flush=True)
    Pv = _rand_matrix(rnd, m, k, espread=4)
    Cv = _rand_matrix(rnd, m, m, espread=4)
    A, B, ex = _exact(Pv, Pv, Cv, False, True, 1)
    Pp = _pack2(Pv)
    got, _ = sdpg.rgemm(PREC, Pp, Pp, _pack2(Cv), False, True, 1, mode, slices, lower=True)
    _check(got, A, B, Cv, ex, 1, bits_lost, lower=True)


@pytest.mark.gpu
@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
@pytest.mark.parametrize("k", [40, 70, 150, 176, 300], ids=["K40", "K70", "K150", "K176", "K300"])
def test_rgemm_small_integers_exact(form, k):
    """K = 40 runs the short-K Ozaki kernel, the others the long-K one (digit sums kept across chunks);
    K = 70, 150, 176 and 300 end in a chunk of 6, 22, 48 and 44 k (zero padding to 64 k)"""
    mode, slices, _ = form
    rnd = random.Random(11 + k)
    m, n = 65, 70
    print(f"case rgemm-int mode={mode} slices={slices} k={k}", flush=True)
    Av = [[Fraction(rnd.randint(-1000, 1000)) for _ in range(k)] for _ in range(m)]
    Bv = [[Fraction(rnd.randint(-1000, 1000)) for _ in range(n)] for _ in range(k)]
    _, _, ex = _exact(Av, Bv, None, False, False, 0)
    got, _ = sdpg.rgemm(PREC, _pack2(Av), _pack2(Bv), None, False, False, 0, mode, slices)
    for i in range(m):
        for j in range(n):
            assert sdpg.words_to_frac(got[i, j], L) == ex[i][j], (i, j)


@pytest.mark.gpu
def test_rgemm_rejects_bad_slices():
    A = _pack2([[Fraction(1)]])
    with pytest.raises(sdpg.SdpgError):
        sdpg.rgemm(PREC, A, A, None, gemm=sdpg.GEMM_OZAKI, slices=7)


@pytest.mark.gpu
@pytest.mark.parametrize("S", [4 * L - 12, 4 * L - 8, 4 * L - 4, 4 * L])
@pytest.mark.parametrize("k", [32, 300], ids=["K32", "K300"])
def test_rgemm_every_slice_count(S, k):
    """every slice count built at 256 bits (28 .. 40), short-K and long-K kernels, within
    K a_i b_j 2^-(8S-12); below S = 4L - 4 the accumulator is narrower than the result's limbs"""
    rnd = random.Random(31 * S + k)
    m, n = 45, 38
    print(f"case rgemm-slices S={S} k={k}", flush=True)
    Av = _rand_matrix(rnd, m, k)
    Bv = _rand_matrix(rnd, n, k)
    Cv = _rand_matrix(rnd, m, n)
    A, B, ex = _exact(Av, Bv, Cv, False, True, 1)
    got, _ = sdpg.rgemm(PREC, _pack2(Av), _pack2(Bv), _pack2(Cv), False, True, 1, sdpg.GEMM_OZAKI, S)
    worst = _check(got, A, B, Cv, ex, 1, 8 * S - 12)
    print(f"  worst error / bound = {worst:.3e}")


def _check_prec(got, Av, Bv, Cv, op, LL, bits_lost):
    """the bar of _check at another limb count"""
    m, n, k = len(Av), len(Bv[0]), len(Bv)
    amax = [max(abs(v) for v in Av[i]) for i in range(m)]
    bmax = [max(abs(Bv[t][j]) for t in range(k)) for j in range(n)]
    ulp = Fraction(1, 1 << (32 * LL - 2))
    for i in range(m):
        for j in range(n):
            ex = sum(Av[i][t] * Bv[t][j] for t in range(k))
            ex = Cv[i][j] - ex if op == 1 else ex
            tol = k * amax[i] * bmax[j] * Fraction(1, 1 << bits_lost) + abs(ex) * ulp + abs(Cv[i][j]) * ulp
            assert abs(sdpg.words_to_frac(got[i, j], LL) - ex) <= tol, (i, j)


@pytest.mark.gpu
@pytest.mark.parametrize("S", [44, 48], ids=["S44", "S48"])
@pytest.mark.parametrize("k", [40, 150, 1326], ids=["K40", "K150", "K1326"])
def test_rgemm_ozaki_320bit(S, k):
    """320 bits (L = 12): the short-K kernel (K = 40) and the long-K kernel at S = 44 and 48 (K = 150
    ends in a 22-k chunk; K = 1326 is the dense rank-2 workload's constraint count, 21 chunks)"""
    prec = 320
    LL = sdpg.limbs_for_bits(prec)
    rnd = random.Random(S + k)
    m, n = 36, 33
    print(f"case rgemm-320 S={S} k={k}", flush=True)

    def rand(r, c):
        out = []
        for _ in range(r):
            row = []
            for _ in range(c):
                M = rnd.getrandbits(32 * LL) | (1 << (32 * LL - 1))
                v = Fraction(M, 1 << (32 * LL)) * Fraction(2) ** rnd.randint(-20, 20)
                row.append(-v if rnd.random() < 0.5 else v)
            out.append(row)
        return out

    Av, Bv, Cv = rand(m, k), rand(k, n), rand(m, n)

    def pack(vals):
        return np.array([[sdpg.frac_to_words(v, LL) for v in row] for row in vals], dtype=np.uint32)

    got, _ = sdpg.rgemm(prec, pack(Av), pack(Bv), pack(Cv), False, False, 1, sdpg.GEMM_OZAKI, S)
    _check_prec(got, Av, Bv, Cv, 1, LL, 8 * S - 12)


@pytest.mark.gpu
def test_rgemm_ozaki_320bit_long_k():
    """At 320 bits (L = 12, S = 44) a K = 1100 product runs on the long-K kernel (18 chunks of 64 k);
    small integers: the exact product sum (no rounding anywhere)."""
    prec, k = 320, 1100
    LL = sdpg.limbs_for_bits(prec)
    rnd = random.Random(23)
    m, n = 33, 35
    print(f"case rgemm-320 ozaki k={k}", flush=True)
    Av = [[Fraction(rnd.randint(-1000, 1000)) for _ in range(k)] for _ in range(m)]
    Bv = [[Fraction(rnd.randint(-1000, 1000)) for _ in range(n)] for _ in range(k)]
    _, _, ex = _exact(Av, Bv, None, False, False, 0)

    def pack(vals):
        return np.array([[sdpg.frac_to_words(v, LL) for v in row] for row in vals], dtype=np.uint32)

    got, _ = sdpg.rgemm(prec, pack(Av), pack(Bv), None, False, False, 0, sdpg.GEMM_OZAKI, 0)
    for i in range(m):
        for j in range(n):
            assert sdpg.words_to_frac(got[i, j], LL) == ex[i][j], (i, j)
