"""Host build of the device arithmetic (mpx.h, compiled for the CPU as libmpxprobe.so) against
exact rational arithmetic: every operation of the limb-float format and the fixed-point
reciprocal square root of the Cholesky pivots.  No GPU needed.

Bars, in units of the result's last place (ulp = 2^(e - 32L)): add/sub/mul <= 0.5 ulp (round to
nearest), div < 1.5 ulp, sqrt/recip < 1 ulp, rsqrt < 2 ulp, rsqrt_fx (fixed point, 2 integer bits, one guard
word) < 2 ulp.
"""
import ctypes
import os
import random
from fractions import Fraction

import pytest

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sdpa-gmp_amd", "lib",
                   "libmpxprobe.so")


def _lib():
    lib = ctypes.CDLL(LIB)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    lib.mpx_probe_op.argtypes = [ctypes.c_int, ctypes.c_int, u32p, u32p, u32p]
    return lib


def _words(x, L):
    if x == 0:
        return [(-(1 << 29)) & 0xFFFFFFFF, 0] + [0] * L
    s = 1 if x < 0 else 0
    x = abs(x)
    e = x.numerator.bit_length() - x.denominator.bit_length()
    while Fraction(2) ** (e - 1) > x:
        e -= 1
    while Fraction(2) ** e <= x:
        e += 1
    M = int(x * Fraction(2) ** (32 * L - e))
    return [e & 0xFFFFFFFF, s] + [(M >> (32 * k)) & 0xFFFFFFFF for k in range(L)]


def _val(w, L):
    M = sum(int(w[2 + k]) << (32 * k) for k in range(L))
    if M == 0:
        return Fraction(0), 0
    e = w[0] if w[0] < (1 << 31) else w[0] - (1 << 32)
    v = Fraction(M) * Fraction(2) ** (e - 32 * L)
    return (-v if w[1] & 1 else v), e


@pytest.mark.parametrize("L", [4, 10, 16])
def test_ops_against_exact(L):
    lib = _lib()
    rnd = random.Random(L)
    A = ctypes.c_uint32 * (L + 2)
    bars = {0: Fraction(1, 2), 1: Fraction(1, 2), 2: Fraction(1, 2), 3: Fraction(3, 2), 4: 1, 5: 1, 6: 2, 7: 2}
    for t in range(300):
        a = Fraction(rnd.getrandbits(32 * L) | 1 << (32 * L - 1), 1 << (32 * L)) * Fraction(2) ** rnd.randint(-90, 90)
        b = Fraction(rnd.getrandbits(32 * L) | 1 << (32 * L - 1), 1 << (32 * L)) * Fraction(2) ** rnd.randint(-90, 90)
        if t % 3 == 0:
            b = -a * (1 + Fraction(rnd.randint(1, 99), 1 << rnd.randint(10, 32 * L - 8)))
        if t % 5 == 1:
            a = -a
        wa, wb = _words(a, L), _words(b, L)
        a, _ = _val(wa, L)
        b, _ = _val(wb, L)
        for op in range(8):
            if op in (4, 6, 7) and a <= 0:
                continue
            r = A()
            assert lib.mpx_probe_op(L, op, A(*wa), A(*wb), r) == 0
            got, e = _val(list(r), L)
            if op == 0:
                ex = a + b
            elif op == 1:
                ex = a - b
            elif op == 2:
                ex = a * b
            elif op == 3:
                ex = a / b
            elif op == 5:
                ex = 1 / a
            else:
                ex = None
            ulp = Fraction(2) ** (e - 32 * L)
            if ex is not None:
                if ex == 0:
                    assert got == 0
                    continue
                assert abs(got - ex) <= bars[op] * ulp, (op, float((got - ex) / ulp))
            elif op == 4:  # sqrt: |got^2 - a| <= 2 ulp(got) * got roughly
                assert abs(got * got - a) <= 2 * bars[op] * ulp * got, (op,)
            else:  # rsqrt: |got^2 a - 1| <= 2 * bar * ulp / got
                assert abs(got * got * a - 1) <= 2 * bars[op] * ulp / got, (op, float((got * got * a - 1) * got / ulp))


@pytest.mark.parametrize("L", [4, 10, 16])
def test_fms_against_exact(L):
    """mpx::fms, a*b - c*d with one rounding (the Bareiss update of the Cholesky diagonal blocks):
    within half an ulp of the exact value plus the products' skipped low columns (below L - 2),
    which weigh below 2 (L - 2) 2^-(32L+30) <= 2^-(32L+24) of the larger product for L <= 18; exact
    zero when the products cancel exactly."""
    lib = _lib()
    lib.mpx_probe_fms.argtypes = [ctypes.c_int] + [ctypes.POINTER(ctypes.c_uint32)] * 5
    rnd = random.Random(100 + L)
    A = ctypes.c_uint32 * (L + 2)

    def rand_mp(spread=40):
        return (Fraction(rnd.getrandbits(32 * L) | 1 << (32 * L - 1), 1 << (32 * L))
                * Fraction(2) ** rnd.randint(-spread, spread) * (-1 if rnd.random() < 0.5 else 1))

    for t in range(400):
        a, b, c, d = rand_mp(), rand_mp(), rand_mp(), rand_mp()
        kind = t % 5
        if kind == 1:  # exact cancellation
            c, d = a, b
        elif kind == 2:  # near cancellation: c d = a b (1 + tiny)
            c = a * (1 + Fraction(rnd.randint(1, 99), 1 << rnd.randint(8, 32 * L - 8)))
            d = b
        elif kind == 3:  # a zero product on either side
            if rnd.random() < 0.5:
                a = Fraction(0)
            else:
                d = Fraction(0)
        w = [_words(x, L) for x in (a, b, c, d)]
        a, b, c, d = (_val(x, L)[0] for x in w)
        r = A()
        assert lib.mpx_probe_fms(L, *(A(*x) for x in w), r) == 0
        got, e = _val(list(r), L)
        ex = a * b - c * d
        if ex == 0:
            assert got == 0, t
            continue
        bar = Fraction(2) ** (e - 32 * L) / 2 + (abs(a * b) + abs(c * d)) * Fraction(2) ** -(32 * L + 24)
        assert abs(got - ex) <= bar, (t, kind, float((got - ex) / Fraction(2) ** (e - 32 * L)))
