"""Host build of the block-column elimination of the blocked Cholesky (chol_fx.h, run serially through
libmpxprobe.so in the order the device kernel follows) against exact rational arithmetic.  No GPU needed.

A block column is the nb x nb diagonal block of B and R rows of the panel below it, with those rows'
diagonal entries.  The factor it returns is checked two ways, exactly:
  - backward: |(L L^T)_rc - B_rc| <= bar sqrt(|B_rr| B_cc) for every entry r >= c of the block column
    (the bar is set by the fixed point's 32L - 4 fraction bits; nb steps of truncated products);
  - the reciprocal pivots: |rdiag_c L_cc - 1| <= 2^-(32L - 8).
Matrices: random well-conditioned ones, and B = S U D U^T S with U unit lower triangular, pivots D over
2^-240 .. 2 and row scales S over 2^-120 .. 2^120 (the regime of late interior-point iterates: the Schur
matrix's pivots spread over many binades and the rows differ in scale).  Failures: B = U D U^T with
d_f = -1 reports info = f (Rpotrf2's first non-positive pivot, mplapack/Rpotrf2.cpp:86-97) for f in the
block, and nothing for f among the panel rows (their pivots belong to later blocks).
"""
import ctypes
import math
import os
import random
from fractions import Fraction

import pytest

from test_mpx_host import LIB, _val, _words

HERE = os.path.dirname(os.path.abspath(__file__))


def _lib():
    lib = ctypes.CDLL(LIB)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    lib.mpx_probe_potrf_col.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, u32p, u32p, u32p, u32p,
                                        ctypes.POINTER(ctypes.c_int)]
    return lib


def _dyadic(rnd, bits, lo=-1, hi=1):
    """a random dyadic rational in [lo, hi) with the given number of bits"""
    return Fraction(lo) + Fraction(rnd.getrandbits(bits), 1 << bits) * (hi - lo)


def ldl_matrix(n, rnd, pivots, scales):
    """B = S U D U^T S (exact Fractions, row-major list of lists)"""
    U = [[Fraction(0)] * n for _ in range(n)]
    for i in range(n):
        U[i][i] = Fraction(1)
        for j in range(i):
            U[i][j] = _dyadic(rnd, 20, -1, 1) / 4
    B = [[Fraction(0)] * n for _ in range(n)]
    for i in range(n):
        for j in range(i + 1):
            s = sum(U[i][k] * pivots[k] * U[j][k] for k in range(j + 1))
            B[i][j] = B[j][i] = s * scales[i] * scales[j]
    return B


def gram_matrix(n, rnd):
    G = [[_dyadic(rnd, 24) for _ in range(n)] for _ in range(n)]
    B = [[sum(G[i][k] * G[j][k] for k in range(n)) + (n if i == j else 0) for j in range(n)] for i in range(n)]
    return B


def run(lib, B, nb, R, L):
    """factor the block column B[:, :nb] (rows 0..nb+R-1) with the panel diagonals B[p][p]"""
    W = L + 2
    arr = ctypes.c_uint32 * ((nb + R) * nb * W)
    A, Lout = arr(), arr()
    for c in range(nb):
        for r in range(c, nb + R):
            A[(r + (nb + R) * c) * W:(r + (nb + R) * c + 1) * W] = _words(B[r][c], L)
    pd = (ctypes.c_uint32 * max(1, R * W))()
    for p in range(R):
        pd[p * W:(p + 1) * W] = _words(B[nb + p][nb + p], L)
    rd = (ctypes.c_uint32 * (nb * W))()
    info = ctypes.c_int(-1)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    cast = lambda x: ctypes.cast(x, u32p)  # noqa: E731
    assert lib.mpx_probe_potrf_col(L, nb, R, cast(A), cast(pd), cast(Lout), cast(rd), ctypes.byref(info)) == 0
    Lm = {}
    if info.value == 0:
        for c in range(nb):
            for r in range(c, nb + R):
                Lm[r, c] = _val(Lout[(r + (nb + R) * c) * W:(r + (nb + R) * c + 1) * W], L)[0]
    rdv = [_val(rd[c * W:(c + 1) * W], L)[0] for c in range(nb)]
    return info.value, Lm, rdv


def check_factor(B, Lm, rdv, nb, R, L):
    worst = Fraction(0)
    for c in range(nb):
        for r in range(c, nb + R):
            s = sum(Lm[r, k] * Lm[c, k] for k in range(c + 1))
            sc = math.sqrt(abs(float(B[r][r])) * float(B[c][c]))
            e = abs(s - B[r][c])
            if e:
                worst = max(worst, Fraction(float(e) / sc) if sc else e)
    rworst = max(abs(rdv[c] * Lm[c, c] - 1) for c in range(nb))
    return float(worst), float(rworst)


def backward_bar(L, nb):
    return 2.0 ** -(32 * L - 4) * 16 * nb


@pytest.mark.parametrize("L", [4, 10])
@pytest.mark.parametrize("nb,R", [(32, 4), (32, 0), (17, 3), (1, 2)])
def test_well_conditioned(L, nb, R):
    lib = _lib()
    rnd = random.Random(1000 * L + nb + R)
    B = gram_matrix(nb + R, rnd)
    info, Lm, rdv = run(lib, B, nb, R, L)
    assert info == 0
    bw, rw = check_factor(B, Lm, rdv, nb, R, L)
    print(f"L={L} nb={nb} R={R}: backward {bw:.3e} (bar {backward_bar(L, nb):.3e}), rdiag {rw:.3e}")
    assert bw <= backward_bar(L, nb)
    assert rw <= 2.0 ** -(32 * L - 8)


@pytest.mark.parametrize("L", [4, 10])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_ill_conditioned_and_scaled(L, seed):
    lib = _lib()
    rnd = random.Random(seed)
    nb, R = 32, 5
    n = nb + R
    pivots = [Fraction(2) ** rnd.randint(-240 if L >= 10 else -60, 1) * (1 + _dyadic(rnd, 16, 0, 1)) for _ in range(n)]
    scales = [Fraction(2) ** rnd.randint(-120, 120) for _ in range(n)]
    B = ldl_matrix(n, rnd, pivots, scales)
    info, Lm, rdv = run(lib, B, nb, R, L)
    assert info == 0
    bw, rw = check_factor(B, Lm, rdv, nb, R, L)
    print(f"L={L} seed={seed}: backward {bw:.3e} (bar {backward_bar(L, nb):.3e}), rdiag {rw:.3e}")
    assert bw <= backward_bar(L, nb)
    assert rw <= 2.0 ** -(32 * L - 8)
    # the pivots themselves: L_cc^2 = d_c s_c^2 up to the cancellation B_cc / (d_c s_c^2) that forming
    # the pivot costs (any backward-stable elimination loses that much)
    for c in range(nb):
        rel = abs(Lm[c, c] ** 2 / (pivots[c] * scales[c] ** 2) - 1)
        bar = 2.0 ** -(32 * L - 16) * nb * float(B[c][c] / (pivots[c] * scales[c] ** 2))
        assert rel <= bar, (c, float(rel), bar)


@pytest.mark.parametrize("fail_at", [1, 2, 9, 32, 33, 36])
def test_failing_pivot(fail_at):
    lib = _lib()
    L = 10
    rnd = random.Random(fail_at)
    nb, R = 32, 4
    n = nb + R
    pivots = [Fraction(1 + rnd.randint(0, 2)) for _ in range(n)]
    pivots[fail_at - 1] = Fraction(-1)
    B = ldl_matrix(n, rnd, pivots, [Fraction(1)] * n)
    info, Lm, rdv = run(lib, B, nb, R, L)
    assert info == (fail_at if fail_at <= nb else 0)
    if info == 0:
        bw, _ = check_factor(B, Lm, rdv, nb, R, L)
        assert bw <= backward_bar(L, nb)


def _fx_words(v, N):
    """canonical limbs of the integer v (29-bit lower limbs, signed top limb) as uint32 words"""
    out = []
    for k in range(N - 1):
        out.append(v & ((1 << 29) - 1))
        v >>= 29
    assert -(1 << 29) <= v < (1 << 29)
    out.append(v & 0xFFFFFFFF)
    return out


def _fx_int(w, N):
    v = 0
    for k in range(N - 1, -1, -1):
        x = w[k] if k < N - 1 else (w[k] - (1 << 32) if w[k] >= (1 << 31) else w[k])
        v = (v << 29) + x
    return v


@pytest.mark.parametrize("L", [4, 8, 10, 12])
def test_fx_update(L):
    """one step's update r = (a b - c d) 2^-s in the fixed point (chol_fx.h fx_upd, the product tier chosen
    from s) against exact integers: within 1 unit of the last bit where partial products are dropped (they
    weigh below 2^-10 of a unit; the result is the floor), exactly the floor with every column kept
    (s < -68); out-of-range results reported."""
    lib = ctypes.CDLL(LIB)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    lib.mpx_probe_fx_upd.argtypes = [ctypes.c_int, u32p, u32p, u32p, u32p, ctypes.c_int, u32p,
                                     ctypes.POINTER(ctypes.c_int)]
    N = lib.mpx_probe_fx_limbs(L)
    F = 29 * N - 3
    rnd = random.Random(L)
    arr = ctypes.c_uint32 * N
    worst = {}
    for it in range(600):
        s = rnd.choice([4, 2, 1, 0, 0, -1, -3, -10, -11, -40, -68, -69, -100, -200])
        a = rnd.randrange(1 << (F - 1), 1 << (F + 1))  # pivot in [1/2, 2)
        b = rnd.randrange(-(1 << (F + 1)), 1 << (F + 1))
        d = rnd.randrange(1 << (F - 1), 1 << (F + 1)) * rnd.choice([-1, 1])
        # c so that a b - c d is about t 2^(F + s) units of 2^-2F with |t| < 4 (|result| < 4)
        t = rnd.randrange(-(1 << (F + 1)), 1 << (F + 1))
        target = a * b - (t << F) * (1 << s) if s >= 0 else a * b - ((t << F) >> -s)
        c = target // d
        if not -(8 << F) < c < (8 << F):
            continue
        exact = a * b - c * d
        ex = exact >> (F + s) if F + s >= 0 else exact << -(F + s)
        out, h = arr(), ctypes.c_int()
        rc = lib.mpx_probe_fx_upd(L, arr(*_fx_words(a, N)), arr(*_fx_words(b, N)), arr(*_fx_words(c, N)),
                                  arr(*_fx_words(d, N)), s, out, ctypes.byref(h))
        if not -(8 << F) <= ex < (8 << F):
            assert rc == 1
            continue
        assert rc == 0, (s, ex)
        got = _fx_int(list(out), N)
        bar = 0 if s < -68 else 1
        assert abs(got - ex) <= bar, (L, s, got - ex)
        key = "exact" if s < -68 else "truncated"
        worst[key] = max(worst.get(key, 0), abs(got - ex))
    print(f"L={L} N={N}: worst |error| in units of 2^-F by tier: {worst}")


def test_block_column_enumerations():
    """the block-column kernel gives each lane one entry at set-up and one item per step by closed-form index
    arithmetic (chol_fx.h entry_code / item_code): every entry of the block column (the block's lower
    triangle, the panel entries, the panel diagonals) exactly once; every item exactly once, panel diagonals
    first, then by non-increasing column, so that the items still active at step j are a prefix"""
    lib = ctypes.CDLL(LIB)
    ip = ctypes.POINTER(ctypes.c_int)
    lib.mpx_probe_col_enum.argtypes = [ctypes.c_int, ctypes.c_int, ip, ip]
    for nb in range(1, 33):
        for R in range(0, 24):
            ne = nb * (nb + 1) // 2 + R * (nb + 1)
            items = (ctypes.c_int * max(1, ne))()
            ents = (ctypes.c_int * max(2, 2 * ne))()
            n = lib.mpx_probe_col_enum(nb, R, items, ents)
            assert n == nb * (nb - 1) // 2 + R * nb
            got = [(items[t] & 0xFFFF, (items[t] >> 16) & 0xFFFF) for t in range(n)]
            want = {(r, c) for c in range(1, nb) for r in range(c, nb + R)} | {(nb + p, 0xFFFF) for p in range(R)}
            assert sorted(got) == sorted(want) and len(set(got)) == n, (nb, R)
            key = [c if c != 0xFFFF else nb for _, c in got]
            assert all(key[t] >= key[t + 1] for t in range(n - 1)), (nb, R)
            for j in range(nb - 1):  # the active items at step j are a prefix
                act = [c > j for _, c in got]
                assert act == sorted(act, reverse=True), (nb, R, j)
            e = [(ents[2 * t], ents[2 * t + 1]) for t in range(ne)]
            want_e = {(r, c) for c in range(nb) for r in range(c, nb + R)} | {(nb + p, -1) for p in range(R)}
            assert sorted(e) == sorted(want_e) and len(set(e)) == ne, (nb, R)
