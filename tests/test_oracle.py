"""CPU tests pinning the oracle (oracle/, a C restatement of the reference solver).

1. Against the reference's own recorded regression results (example1.result and
   SDPLIB/data/*.result.xz, copied under tests/golden/): same phase, same iteration count,
   objectives to the 16 printed digits, relative gap to the printed digits, and the
   per-iteration mu / objective columns of the log.
2. Arithmetic (ompf: GMP mpf semantics) against exact rational arithmetic.
3. The Newton step on a workload-0 iterate against closed forms computed exactly with
   Python integers: B_ij = Tr(A_i X A_j Z^-1) (F1 formula), L L^T = B, B dy = g,
   dZ = -sum dy_i A_i (+Rd), dX = sym(beta mu Z^-1 - X - X dZ Z^-1).
"""
import json
import os
import random
import re
import tempfile
from decimal import Decimal
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

FAST = ["example1", "control1", "truss1", "truss4", "hinf1", "hinf3", "infp1", "infd1", "theta1", "control2"]


def _gold():
    with open(os.path.join(GOLD, "sdplib_golden.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", FAST)
def test_oracle_matches_reference_results(name, sdplib_file):
    g = _gold()[name]
    log = tempfile.mktemp(suffix=".log")
    r = oracle_lib.solve(sdplib_file(name), None, 0, log)
    assert r.phase_name.decode().strip() == g["phase"]
    assert r.iterations == g["iterations"]
    assert Decimal(r.objP.decode()).quantize(Decimal(g["objP"])) == Decimal(g["objP"]) or \
        abs(Decimal(r.objP.decode()) - Decimal(g["objP"])) <= abs(Decimal(g["objP"])) * Decimal("1e-16")
    assert abs(Decimal(r.objD.decode()) - Decimal(g["objD"])) <= abs(Decimal(g["objD"])) * Decimal("1e-16") + \
        Decimal("1e-300")
    # relative gap: printed with 17 digits; agree to 8 (it is a rounding-level quantity for pdOPT)
    if g["relgap"]:
        a, b = Decimal(r.relgap_s.decode()), Decimal(g["relgap"])
        assert abs(a - b) <= abs(b) * Decimal("1e-8")
    # per-iteration log: mu (2 digits) and the two objectives (3 digits) agree to the printed
    # rounding (the last digit may differ by one at a rounding boundary)
    with open(log) as f:
        mine = [ln for ln in f if re.match(r"^\s*\d+ \d\.\de[+-]\d\d ", ln)]
    assert len(mine) == len(g["log"])
    for a, b in zip(mine, g["log"]):
        pa, pb = a.split(), b.split()
        for col, rel in ((1, 0.051), (4, 0.0051), (5, 0.0051)):
            va, vb = float(pa[col]), float(pb[col])
            assert abs(va - vb) <= rel * max(abs(vb), 1e-300) + 0.011 * (abs(vb) < 0.1), (a, b)


def _words(x, L):
    """Fraction -> (sign, exp, limbs) of the oracle interchange format (truncating)."""
    if x == 0:
        return 0, 0, [0] * L
    s = -1 if x < 0 else 1
    x = abs(x)
    e = x.numerator.bit_length() - x.denominator.bit_length() + 1
    while Fraction(2) ** (e - 1) > x:
        e -= 1
    while Fraction(2) ** e <= x:
        e += 1
    M = int(x * Fraction(2) ** (32 * L - e))
    return s, e, [(M >> (32 * k)) & 0xFFFFFFFF for k in range(L)]


def _val(s, e, limbs):
    L = len(limbs)
    M = sum(int(v) << (32 * k) for k, v in enumerate(limbs))
    return Fraction(s) * Fraction(M) * Fraction(2) ** (e - 32 * L) if M else Fraction(0)


@pytest.mark.parametrize("bits", [200, 256, 512])
def test_oracle_arithmetic_against_exact(bits):
    import ctypes

    lib = oracle_lib.lib()
    prec_limbs = (max(53, bits) + 127) // 64
    Lx = 2 * (prec_limbs + 1)  # exact interchange width of a result
    rnd = random.Random(bits)
    U32 = ctypes.c_uint32
    for t in range(400):
        a = Fraction(rnd.getrandbits(bits), 1) * Fraction(2) ** rnd.randint(-300, 300) * rnd.choice([-1, 1])
        b = Fraction(rnd.getrandbits(bits), 1) * Fraction(2) ** rnd.randint(-300, 300) * rnd.choice([-1, 1])
        if t % 4 == 0:
            b = -a * (1 + Fraction(rnd.randint(1, 1000), 2 ** (bits - 20)))  # near cancellation
        # inputs carry 64*prec bits so that they fit prec+1 limbs at any bit alignment
        sa, ea, la = _words(a, 2 * prec_limbs)
        sb, eb, lb = _words(b, 2 * prec_limbs)
        a, b = _val(sa, ea, la), _val(sb, eb, lb)
        la, lb = [0, 0] + la, [0, 0] + lb
        for op in range(5):
            if op == 3 and b == 0:
                continue
            if op == 4 and a < 0:
                continue
            rs, re_, rl = ctypes.c_int32(), ctypes.c_int32(), (U32 * Lx)()
            rc = lib.orc_arith(bits, op, Lx, sa, ea, (U32 * Lx)(*la), sb, eb, (U32 * Lx)(*lb), ctypes.byref(rs),
                               ctypes.byref(re_), rl)
            assert rc == 0
            got = _val(rs.value, re_.value, list(rl))
            ex = [a + b, a - b, a * b, a / b if b else 0, None][op]
            if op == 4:
                if a == 0:
                    continue
                err = abs(got * got - a) / abs(a) / 2
            else:
                if ex == 0:
                    continue
                err = abs(got - ex) / abs(ex)
            # GMP-style truncation: relative error below 2^-(64*prec - 64), i.e. the requested precision
            assert err < Fraction(2) ** -(64 * prec_limbs - 64 - 2), (op, float(err))
            # products truncate toward zero (a sum may not: an operand whose exponent lies a
            # full precision below the other's is dropped; a quotient may not: the divisor is
            # truncated first)
            if op == 2:
                assert abs(got) <= abs(ex), op


def test_oracle_decimal_conversion_exact():
    import ctypes

    lib = oracle_lib.lib()
    for text in ["3.4495156687107737525366246700286865234375", "-2.589385553948886808939278125762939453125",
                 "1.0E4", "-1e-30", "0.1", "48", "123456789012345678901234567890"]:
        s, e, L = ctypes.c_int32(), ctypes.c_int32(), 12
        limbs = (ctypes.c_uint32 * L)()
        assert lib.orc_parse(256, text.encode(), L, ctypes.byref(s), ctypes.byref(e), limbs) == 0
        got = _val(s.value, e.value, list(limbs))
        ex = Fraction(Decimal(text))
        assert abs(got - ex) <= abs(ex) * Fraction(2) ** -318  # exact for dyadic text, 5+1 limbs otherwise


# ---------------------------------------------------------------- Newton step, exact closed forms

def _frac_arr(packed):
    L = packed.shape[1] - 2
    out = []
    for row in packed:
        M = 0
        for k in range(L - 1, -1, -1):
            M = (M << 32) | int(row[2 + k])
        if M == 0:
            out.append(Fraction(0))
            continue
        e = int(np.int32(np.uint32(row[0])))
        v = Fraction(M) * Fraction(2) ** (e - 32 * L)
        out.append(-v if int(row[1]) & 1 else v)
    return out


def _read_problem(path):
    """Dense symmetric A_i (exact Fractions) of a single-block .dat-s with n x n blocks."""
    with open(path) as f:
        lines = [ln for ln in f if not ln.startswith('"')]
    m = int(lines[0].split()[0])
    n = int(lines[2].split()[0])
    b = [Fraction(Decimal(t)) for t in lines[3].split()]
    A = [[[Fraction(0)] * n for _ in range(n)] for _ in range(m + 1)]
    for ln in lines[4:]:
        p = ln.split()
        k, _, i, j = (int(x) for x in p[:4])
        v = Fraction(Decimal(p[4]))
        if k == 0:
            v = -v
        A[k][i - 1][j - 1] = v
        A[k][j - 1][i - 1] = v
    return m, n, b, A[0], A[1:]


def _mm(P, Q):
    n = len(P)
    return [[sum(P[i][k] * Q[k][j] for k in range(n)) for j in range(n)] for i in range(n)]


def _mat(vals, n):
    return [[vals[i + n * j] for j in range(n)] for i in range(n)]


def test_oracle_newton_step_closed_forms(workload):
    """Oracle 256-bit Newton step on workload 0, iteration 1, against exact rational algebra."""
    path = workload(0)
    m, n, b, C, A = _read_problem(path)
    fx = dict(np.load(os.path.join(GOLD, "newton_w0_it001.npz")))
    X = _mat(_frac_arr(fx["in_X"]), n)
    invZ = _mat(_frac_arr(fx["in_invZ"]), n)
    # B_ij = Tr(A_i X A_j Z^-1): exact with G_j = X A_j Z^-1
    G = [_mm(_mm(X, A[j]), invZ) for j in range(m)]
    Bex = [[sum(A[i][p][q] * G[j][q][p] for p in range(n) for q in range(n)) for j in range(m)] for i in range(m)]
    Bref = _frac_arr(fx["ref_B"])
    scale = max(abs(v) for row in Bex for v in row)
    err = max(abs(Bref[i + m * j] - Bex[i][j]) for i in range(m) for j in range(m))
    assert err / scale < Fraction(1, 10 ** 120)  # the 512-bit reference is exact to its precision
    # symmetry of the exact definition
    assert all(Bex[i][j] == Bex[j][i] for i in range(m) for j in range(m))
    # Cholesky: L L^T = B
    Lf = fx["ref_L"]
    Lv = _frac_arr(Lf)
    Lm = [[Fraction(0)] * m for _ in range(m)]
    t = 0
    for j in range(m):
        for i in range(m):
            if i >= j:
                Lm[i][j] = Lv[t]
                t += 1
    LLt = [[sum(Lm[i][k] * Lm[j][k] for k in range(m)) for j in range(m)] for i in range(m)]
    err = max(abs(LLt[i][j] - Bex[i][j]) for i in range(m) for j in range(m))
    assert err / scale < Fraction(1, 10 ** 120)
    # B dy = g
    dy = _frac_arr(fx["ref_dy"])
    g = _frac_arr(fx["ref_g"])
    res = max(abs(sum(Bex[i][j] * dy[j] for j in range(m)) - g[i]) for i in range(m))
    assert res / max(abs(v) for v in g) < Fraction(1, 10 ** 110)
    # dZ = -sum dy_i A_i (+ Rd when the phase is infeasible)
    phase = int(fx["phase"][0])
    Rd = _mat(_frac_arr(fx["ref_Rd"]), n)
    dZ = _mat(_frac_arr(fx["ref_dZ"]), n)
    for p in range(n):
        for q in range(n):
            ex = -sum(dy[i] * A[i][p][q] for i in range(m)) + (Rd[p][q] if phase in (0, 1) else 0)
            assert abs(dZ[p][q] - ex) <= Fraction(1, 10 ** 120) * (1 + abs(ex))
    # dX = sym(beta mu Z^-1 - X - X dZ Z^-1)
    mu = _frac_arr(fx["in_mu"])[0]
    beta = _frac_arr(fx["in_beta"])[0]
    XdZZ = _mm(_mm(X, dZ), invZ)
    dX = _mat(_frac_arr(fx["ref_dX"]), n)
    for p in range(n):
        for q in range(n):
            u = beta * mu * invZ[p][q] - X[p][q] - XdZZ[p][q]
            v = beta * mu * invZ[q][p] - X[q][p] - XdZZ[q][p]
            ex = (u + v) / 2
            assert abs(dX[p][q] - ex) <= Fraction(1, 10 ** 110) * (1 + abs(ex))


def test_oracle_512_converges_on_workload0(workload):
    """The 512-bit oracle run that the GPU's 256-bit solve is compared with (solves.json)."""
    with open(os.path.join(GOLD, "solves.json")) as f:
        g = json.load(f)["w0"]["oracle512"]
    param = os.path.join(os.path.dirname(GOLD), "..", "sdpa-gmp_amd", "data", "param_256_gap1e-60.sdpa")
    r = oracle_lib.solve(workload(0), param, 512, None)
    assert r.phase_name.decode().strip() == g["phase"] == "pdOPT"
    assert r.iterations == g["iterations"]
    assert r.objP70.decode() == g["objP"] and r.objD70.decode() == g["objD"]
