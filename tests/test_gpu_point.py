"""Device-resident point handling against exact arithmetic on the oracle's iterates.

- sdpg_factor_point (Solutions::computeInverse, sdpa_dataset.cpp:96-118 -> Jal::getInvChol /
  getInvCholAndInv, sdpa_jordan.cpp:123-171): invCholX X invCholX^T = I, invCholZ Z invCholZ^T = I,
  invZ Z = I and lower-triangular factors, checked in exact integer arithmetic.
- sdpg_update (Solutions::update, sdpa_dataset.cpp:120-133): X + alpha dX etc. element by element.
- sdpg_inner: the scalar inner products of StepLength / DirectionParameter / SolveInfo against exact
  sums of the downloaded arrays.
- sdpg_min_eigen (Jal::getMinEigen, sdpa_jordan.cpp:55-88): blocks <= 20 (the reference's Rsyev path)
  to 1e-12 relative against numpy's eigvalsh of the exact product rounded to float64; blocks > 20
  (Lanczos, sdpa_linear.cpp:27-124) within the reference's own stopping tolerance.
"""
import os
from fractions import Fraction

import numpy as np
import pytest

import sdpg
from test_gpu_newton import GOLD, run_step

CASES = ["newton_w0_it001.npz", "newton_w0_it005.npz", "newton_w1_it010.npz"]


def _blocks(ctx, path):
    with open(path) as f:
        lines = [ln for ln in f if not ln.startswith('"') and not ln.startswith("*")]
    nb = int(lines[1].split()[0])
    sizes = [int(x) for x in lines[2].replace("=", " ").replace(",", " ").replace("{", " ").replace("}", " ")
             .replace("(", " ").replace(")", " ").split() if x.lstrip("-").isdigit()][:nb]
    return [n for n in sizes if n >= 2]


def _ints(packed, L):
    """packed mp values -> (python ints, common exponent s) with value = int * 2^s"""
    vals = sdpg.unpack(packed, L)
    s = 0
    for v in vals:
        if v:
            d = v.denominator.bit_length() - 1
            s = min(s, -d)
    return [int(v * (Fraction(2) ** -s)) for v in vals], s


def _mat(ints, o, n):
    return np.array([[ints[o + i + n * j] for j in range(n)] for i in range(n)], dtype=object)


def _load(ctx, fx):
    L = ctx.L
    rp = lambda k: sdpg.repack(fx["in_" + k], fx["in_" + k].shape[1] - 2, L)  # noqa: E731
    ctx.set_point(rp("X"), rp("Z"), rp("invZ"), rp("y"))
    return rp


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_factor_point_inverses(case, workload):
    fx = dict(np.load(os.path.join(GOLD, case)))
    wk = int(case.split("_")[1][1:])
    print(f"case factor_point {case}", flush=True)
    ctx = sdpg.Context(workload(wk), 256)
    try:
        _load(ctx, fx)
        assert ctx.factor_point() == 0
        L = ctx.L
        got = {k: _ints(ctx.get(sel), L) for k, sel in (("X", sdpg.SDPG_X), ("Z", sdpg.SDPG_Z),
                                                         ("icX", sdpg.SDPG_INVCHOLX), ("icZ", sdpg.SDPG_INVCHOLZ),
                                                         ("iZ", sdpg.SDPG_INVZ))}
        sizes = _blocks(ctx, workload(wk))
    finally:
        ctx.close()
    o = 0
    for n in sizes:
        M = {k: _mat(v[0], o, n) for k, v in got.items()}
        S = {k: v[1] for k, v in got.items()}
        for a, x in (("icX", "X"), ("icZ", "Z")):
            Lm = M[a]
            assert all(Lm[i, j] == 0 for i in range(n) for j in range(i + 1, n)), f"{a} not lower triangular"
            R = Lm.dot(M[x]).dot(Lm.T)
            sc = 2 * S[a] + S[x]
            one = 1 << -sc
            err = max(abs(int(R[i, j]) - (one if i == j else 0)) for i in range(n) for j in range(n))
            rel = Fraction(err, one)
            print(f"  block {n}: |{a} {x} {a}^T - I| = {float(rel):.3e}")
            assert rel < Fraction(1, 10 ** 60)
        R = M["iZ"].dot(M["Z"])
        sc = S["iZ"] + S["Z"]
        one = 1 << -sc
        err = max(abs(int(R[i, j]) - (one if i == j else 0)) for i in range(n) for j in range(n))
        print(f"  block {n}: |invZ Z - I| = {float(Fraction(err, one)):.3e}")
        assert Fraction(err, one) < Fraction(1, 10 ** 60)
        o += n * n


def _step(ctx, fx):
    out = run_step(ctx, fx)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_inner_products_and_update(case, workload):
    fx = dict(np.load(os.path.join(GOLD, case)))
    wk = int(case.split("_")[1][1:])
    print(f"case inner/update {case}", flush=True)
    ctx = sdpg.Context(workload(wk), 256)
    try:
        L = ctx.L
        out = _step(ctx, fx)
        fr = lambda sel: sdpg.unpack(ctx.get(sel), L)  # noqa: E731
        X, Z, dX, dZ, y, dy = (fr(s) for s in (sdpg.SDPG_X, sdpg.SDPG_Z, sdpg.SDPG_DX, sdpg.SDPG_DZ, sdpg.SDPG_Y,
                                                sdpg.SDPG_DY))
        b = None
        res = ctx.residuals()
        objP = sdpg.words_to_frac(res[2], L)
        ip = lambda w: sdpg.words_to_frac(ctx.inner(w), L)  # noqa: E731
        ex = {
            sdpg.SDPG_IP_XZ: (X, Z), sdpg.SDPG_IP_XDZ: (X, dZ), sdpg.SDPG_IP_DXZ: (dX, Z), sdpg.SDPG_IP_DXDZ: (dX, dZ),
        }
        for w, (a, c) in ex.items():
            e = sum(p * q for p, q in zip(a, c))
            mag = sum(abs(p * q) for p, q in zip(a, c))
            assert abs(ip(w) - e) <= mag * Fraction(1, 2 ** 300) + Fraction(1, 10 ** 300), w
        assert ip(sdpg.SDPG_IP_CX) == objP
        sizes = _blocks(ctx, workload(wk))
        diag = []
        o = 0
        for n in sizes:
            diag += [o + i + n * i for i in range(n)]
            o += n * n
        diag += list(range(o, len(X)))
        for w, M in ((sdpg.SDPG_IP_TRX, X), (sdpg.SDPG_IP_TRZ, Z)):
            e = sum(M[i] for i in diag)
            assert abs(ip(w) - e) <= sum(abs(M[i]) for i in diag) * Fraction(1, 2 ** 300), w
        # b . y and b . dy: b recovered from Rp + A.X is not needed: b . y at y and y + dy
        by0 = ip(sdpg.SDPG_IP_BY)
        bdy = ip(sdpg.SDPG_IP_BDY)
        half = sdpg.pack([Fraction(1, 2)], L)[0]
        ctx.update(half, half)
        X2, Z2, y2 = (fr(s) for s in (sdpg.SDPG_X, sdpg.SDPG_Z, sdpg.SDPG_Y))
        for a0, d, a1 in ((X, dX, X2), (Z, dZ, Z2), (y, dy, y2)):
            for p, q, r in zip(a0, d, a1):
                e = p + q / 2
                assert abs(r - e) <= (abs(p) + abs(q)) * Fraction(1, 2 ** 315), (p, q, r)
        by1 = ip(sdpg.SDPG_IP_BY)
        lin = by0 + bdy / 2
        assert abs(by1 - lin) <= (abs(by0) + abs(bdy)) * Fraction(1, 2 ** 290) + Fraction(1, 10 ** 300)
        del out, b
    finally:
        ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_min_eigen(case, workload):
    fx = dict(np.load(os.path.join(GOLD, case)))
    wk = int(case.split("_")[1][1:])
    print(f"case min_eigen {case}", flush=True)
    ctx = sdpg.Context(workload(wk), 256)
    try:
        L = ctx.L
        _step(ctx, fx)
        assert ctx.factor_point() == 0
        sizes = _blocks(ctx, workload(wk))
        for which, (lsel, dsel) in enumerate(((sdpg.SDPG_INVCHOLX, sdpg.SDPG_DX), (sdpg.SDPG_INVCHOLZ, sdpg.SDPG_DZ))):
            Lv = np.array([float(v) for v in sdpg.unpack(ctx.get(lsel), L)])
            Dv = np.array([float(v) for v in sdpg.unpack(ctx.get(dsel), L)])
            v = float(sdpg.words_to_frac(ctx.min_eigen(which), L))
            o = 0
            best = None
            small_only = True
            for n in sizes:
                Lm = Lv[o:o + n * n].reshape(n, n, order="F")
                Dm = Dv[o:o + n * n].reshape(n, n, order="F")
                lam = np.linalg.eigvalsh(Lm @ Dm @ Lm.T).min()
                best = lam if best is None else min(best, lam)
                small_only = small_only and n <= 20
                o += n * n
            print(f"  which={which}: device {v:.12e}  numpy {best:.12e}")
            if small_only:
                assert abs(v - best) <= 1e-10 * abs(best) + 1e-12
            else:
                assert abs(v - best) <= 0.05 * abs(best) + 1e-3
    finally:
        ctx.close()
