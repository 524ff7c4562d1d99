"""Newton-step parity on the GPU: the engine's Rp, Rd, g, B, L, dy, dZ, dX on the oracle's iterates.

Each step is checked two ways.

Backward / consistency checks, in exact integer arithmetic on the device's own values, with bars set
by the device's arithmetic and independent of the conditioning of B (a slicing or carry slip that
costs 30 bits anywhere fails them at every iterate):
  - B (sdpa_newton.cpp:1007-1096) against the 512-bit oracle's B, entry by entry, relative to
    ||A_i||_1 ||A_j||_1 ||X||_max ||inv(Z)||_max;
  - the factor: D (L L^T - B) D with D = diag(B)^(-1/2) rounded to powers of two (all entries for
    workload 0, two random sign vectors for workload 1);
  - the solve: D (B dy - g), relative to m |D^-1 dy|;
  - dZ = -sum dy_i A_i (+ Rd in the infeasible phases) recomputed exactly from the device's dy
    (sdpa_newton.cpp:1346-1356);
  - dX = sym(beta mu inv(Z) - X - X dZ inv(Z)) recomputed exactly from the device's dZ
    (sdpa_newton.cpp:1358-1373, 821-838).
Forward checks against the 512-bit oracle on the same iterate: Rp, Rd absolutely; g relative to the
size of its terms; L, dy, dZ, dX normwise within max(2^-(P-24), err256 2^-(P-288)), where err256 is
the GMP-256 oracle's own normwise error on that iterate.  Forward errors grow with the condition of
B, and err256 measures it: GMP at a 256-bit request carries five 64-bit limbs with truncation, and
its forward error on these iterates behaves like that of about 280 bits.  Measured on the fixtures,
the plain form's forward error is 2^-42 to 2^-52 of err256 and the Ozaki form's 2^-6.6 to 2^-17, so
the bar leaves at least 2^10 on every iterate.

P is the device precision of the products: 32L bits for the plain limb form; 8S - 4 for the Ozaki
form with S slices (block fixed point of 8S - 2 fraction bits per row / column, sdpg_set_gemm).
Every reference key must be present in the fixture.
"""
import glob
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import exact
import sdpg

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(GOLD, "newton_w*_it*.npz")))
REF_KEYS = ("Rp", "Rd", "g", "B", "L", "dy", "dX", "dZ")


def normwise(got, ref):
    scale = max(abs(x) for x in ref)
    if scale == 0:
        return float(max(abs(x) for x in got))
    return float(max(abs(a - b) for a, b in zip(got, ref)) / scale)


def product_bits(gemm, L, slices=0):
    """precision in bits of the device's products for a GEMM form"""
    if gemm == sdpg.GEMM_OZAKI:
        S = slices or 4 * L - 4
        return 8 * S - 4
    return 32 * L


def forward_bar(P, err256):
    """bar of the forward error against the 512-bit oracle (see the module docstring)"""
    return max(2.0 ** -(P - 24), err256 * 2.0 ** -(P - 288))


def device_inputs(fx, L):
    rp = lambda k: sdpg.repack(fx["in_" + k], fx["in_" + k].shape[1] - 2, L)  # noqa: E731
    return {k: rp(k) for k in ("X", "Z", "invZ", "y", "mu", "beta")}


def run_step(ctx, fx, inp=None):
    inp = inp or device_inputs(fx, ctx.L)
    ctx.set_point(inp["X"], inp["Z"], inp["invZ"], inp["y"])
    ctx.residuals()
    phase = int(fx["phase"][0])
    ctx.gvec(False, inp["mu"][0], inp["beta"][0], phase)
    out = {"Rp": ctx.get(sdpg.SDPG_RP), "Rd": ctx.get(sdpg.SDPG_RD), "g": ctx.get(sdpg.SDPG_GVEC)}
    ctx.schur()
    out["B"] = ctx.get(sdpg.SDPG_B)
    info = ctx.potrf()
    assert info == 0, f"Cholesky failed (info={info})"
    out["L"] = ctx.get(sdpg.SDPG_L)
    ctx.potrs()
    ctx.dz(phase)
    ctx.dx()
    out["dy"] = ctx.get(sdpg.SDPG_DY)
    out["dZ"] = ctx.get(sdpg.SDPG_DZ)
    out["dX"] = ctx.get(sdpg.SDPG_DX)
    return out


def _lower_positions(m):
    return np.array([i + m * j for j in range(m) for i in range(j, m)], dtype=np.int64)


def _pow2_scale(bdiag):
    """D_i = 2^-floor(log2(B_ii) / 2) as exponents k_i (D_i = 2^-k_i)"""
    return np.array([int(math.floor(math.log2(v) / 2)) if v > 0 else 0 for v in bdiag], dtype=np.int64)


def _scaled(vec_ints, E, k):
    """float64 of 2^-k_i v_i 2^E"""
    return np.array([float(Fraction(int(v)) * Fraction(2) ** (E - int(ki))) if v else 0.0
                     for v, ki in zip(vec_ints, k)])


class Checks:
    def __init__(self, name):
        self.name, self.fail = name, []

    def le(self, key, err, bar, note=""):
        ok = err <= bar
        print(f"  {key:<14} {err:10.3e} (bar {bar:9.3e}, log2 ratio {math.log2(max(err, 1e-300) / bar):7.1f}) "
              f"{'ok' if ok else 'FAIL'} {note}", flush=True)
        if not ok:
            self.fail.append(f"{key}: {err:.3e} > {bar:.3e}")


@pytest.mark.gpu
@pytest.mark.parametrize("gemm", [sdpg.GEMM_PLAIN, sdpg.GEMM_OZAKI], ids=["plain", "ozaki"])
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[7:-4] for p in FIXTURES])
def test_newton_step_matches_oracle(path, gemm, workload):
    name = os.path.basename(path)
    wk = int(name.split("_")[1][1:])
    fx = dict(np.load(path))
    missing = [k for k in REF_KEYS if "ref_" + k not in fx]
    assert not missing, f"fixture {name} lacks reference outputs {missing}"
    print(f"case {name} gemm={gemm}", flush=True)
    ctx = sdpg.Context(workload(wk), 256, gemm=gemm)
    try:
        L = ctx.L
        inp = device_inputs(fx, L)
        out = run_step(ctx, fx, inp)
    finally:
        ctx.close()
    P = product_bits(gemm, L)
    Lr = fx["ref_dy"].shape[1] - 2
    m, sizes, n_lp, bvec, Cm, Am = exact.read_problem(workload(wk))
    assert n_lp == 0
    nmax = max(sizes)
    phase = int(fx["phase"][0])
    chk = Checks(name)

    # float views for the bounds
    f = lambda a: exact.to_float(a, L)  # noqa: E731
    Xb, iZb = exact.block_views(f(inp["X"]), sizes), exact.block_views(f(inp["invZ"]), sizes)
    Zb = exact.block_views(f(inp["Z"]), sizes)
    nX = max(np.abs(x).max() for x in Xb)
    niZ = max(np.abs(x).max() for x in iZb)
    nZ = max(np.abs(x).max() for x in Zb)
    A1 = np.zeros((m, len(sizes)))  # ||A_i^l||_1 (both triangles)
    Amax = np.zeros(m)
    for i, ents in enumerate(Am):
        for blk, r, c, v in ents:
            A1[i, blk] += abs(float(v)) * (1 if r == c else 2)
            Amax[i] = max(Amax[i], abs(float(v)))
    Cmax = max([abs(float(v)) for _, _, _, v in Cm] + [0.0])
    yf = f(inp["y"])
    bm = float(Fraction(sdpg.unpack(inp["beta"], L)[0]) * Fraction(sdpg.unpack(inp["mu"], L)[0]))

    # residuals (plain limb arithmetic in every form)
    res_scale = 1 + nX + nZ + Cmax + float(np.sum(np.abs(yf) * Amax)) + max(abs(float(x)) for x in bvec)
    for key in ("Rp", "Rd"):
        got, ref = sdpg.unpack(out[key], L), sdpg.unpack(fx["ref_" + key], Lr)
        err = float(max(abs(a - b) for a, b in zip(got, ref)))
        chk.le(key, err, 2.0 ** -(32 * L - 24) * res_scale * nmax, "absolute")
    Rdf = f(out["Rd"])
    nRd = float(np.abs(Rdf).max())

    # g: relative to the size of its terms
    infeas = phase in (sdpg.SDPG_NOINFO, sdpg.SDPG_PFEAS)  # the X Rd inv(Z) and Rd terms enter
    dls1 = bm * niZ + nX + (nmax * nX * nRd * niZ if infeas else 0.0)
    g_scale = float(A1.sum(axis=1).max()) * dls1 + float(np.abs(f(out["Rp"])).max())
    g_err = float(max(abs(a - b) for a, b in zip(sdpg.unpack(out["g"], L), sdpg.unpack(fx["ref_g"], Lr))))
    chk.le("g", g_err, 2.0 ** -(P - 16) * nmax * nmax * g_scale)

    # B against the oracle, entrywise relative to ||A_i||_1 ||A_j||_1 ||X|| ||inv Z||
    pos = fx.get("ref_B_idx", np.arange(m * m))
    Bref = exact.to_ints(fx["ref_B"], Lr)
    Bdev_all, EB = exact.to_ints(out["B"], L)
    Bdev = Bdev_all[pos]
    d = exact.diff_float(Bdev, EB, Bref[0], Bref[1])
    ii, jj = pos % m, pos // m
    scale = np.array([sum(A1[i, l] * A1[j, l] for l in range(len(sizes))) for i, j in zip(ii, jj)]) * nX * niZ
    scale = np.maximum(scale, 1e-300)
    chk.le("B", float(np.max(np.abs(d) / scale)), 2.0 ** -(P - 16) * nmax * nmax, f"{len(pos)} entries")

    # factor: backward error D (L L^T - B) D
    Bf = exact.ldexp_obj(Bdev_all, EB).reshape(m, m, order="F")
    kD = _pow2_scale(np.diag(Bf))
    Lint, EL = exact.to_ints(out["L"], L)
    Lmat = Lint.reshape(m, m, order="F")
    Lmat = np.where(np.tril(np.ones((m, m), dtype=bool)), Lmat, 0).astype(object)
    Bmat = Bdev_all.reshape(m, m, order="F")
    if m <= 64:
        R = Lmat.dot(Lmat.T)  # at 2 EL
        Ra, Bb, E = exact.align(R.reshape(-1), 2 * EL, Bmat.reshape(-1), EB)
        Rm = (Ra - Bb).reshape(m, m)
        err = max(abs(float(Fraction(int(Rm[i, j])) * Fraction(2) ** (E - int(kD[i]) - int(kD[j]))))
                  for i in range(m) for j in range(m))
        chk.le("LL^T-B", err, 2.0 ** -(P - 16) * m, "all entries, scaled")
    else:
        rng = np.random.default_rng(7)
        err = 0.0
        for _ in range(2):
            w = rng.choice([-1, 1], size=m)
            # v = D w as integers at exponent -kmax
            kmax = int(kD.max())
            v = np.array([int(w[i]) << (kmax - int(kD[i])) for i in range(m)], dtype=object)
            t = Lmat.T.dot(v)
            r1 = Lmat.dot(t)  # exponent 2 EL - kmax
            r2 = Bmat.dot(v)  # exponent EB - kmax
            a, b, E = exact.align(r1, 2 * EL - kmax, r2, EB - kmax)
            err = max(err, float(np.max(np.abs(_scaled(a - b, E, kD)))))
        chk.le("LL^T-B", err, 2.0 ** -(P - 16) * m * m, "two sign vectors, scaled")

    # solve: D (B dy - g)
    dyi, Edy = exact.to_ints(out["dy"], L)
    gi, Eg = exact.to_ints(out["g"], L)
    r = Bmat.dot(dyi)
    a, b, E = exact.align(r, EB + Edy, gi, Eg)
    rs = np.max(np.abs(_scaled(a - b, E, kD)))
    dyf = exact.ldexp_obj(dyi, Edy)
    chk.le("B dy - g", float(rs), 2.0 ** -(P - 16) * m * m * float(np.max(np.abs(dyf * np.ldexp(1.0, kD)))),
           "scaled")

    # dZ recomputed exactly from the device's dy (plain limb arithmetic)
    dy = sdpg.unpack(out["dy"], L)
    Rd = sdpg.unpack(out["Rd"], L)
    dZ = sdpg.unpack(out["dZ"], L)
    offs = np.cumsum([0] + [n * n for n in sizes])
    exp_dZ = [Fraction(0)] * len(dZ)
    size_dZ = [0.0] * len(dZ)
    for i, ents in enumerate(Am):
        for blk, r_, c_, v in ents:
            n = sizes[blk]
            for (p, q) in {(r_, c_), (c_, r_)}:
                k = int(offs[blk]) + p + n * q
                exp_dZ[k] -= dy[i] * v
                size_dZ[k] += abs(float(dy[i]) * float(v))
    if infeas:
        for k in range(len(dZ)):
            exp_dZ[k] += Rd[k]
            size_dZ[k] += abs(float(Rd[k]))
    err = max(float(abs(a - b)) / (s if s > 0 else 1.0) for a, b, s in zip(dZ, exp_dZ, size_dZ))
    chk.le("dZ(dy)", err, 2.0 ** -(32 * L - 8), "entrywise relative to sum |dy_i A_i|")

    # dX recomputed exactly from the device's dZ: sym(beta mu invZ - X - X dZ invZ)
    Xi, EX = exact.to_ints(inp["X"], L)
    iZi, EiZ = exact.to_ints(inp["invZ"], L)
    dZi, EdZ = exact.to_ints(out["dZ"], L)
    dXd = sdpg.unpack(out["dX"], L)
    bmF = Fraction(sdpg.unpack(inp["beta"], L)[0]) * Fraction(sdpg.unpack(inp["mu"], L)[0])
    err = 0.0
    for blk, n in enumerate(sizes):
        o = int(offs[blk])
        Xm = Xi[o:o + n * n].reshape(n, n, order="F")
        iZm = iZi[o:o + n * n].reshape(n, n, order="F")
        dZm = dZi[o:o + n * n].reshape(n, n, order="F")
        T = Xm.dot(dZm).dot(iZm)  # exponent EX + EdZ + EiZ
        nT = max(np.abs(exact.ldexp_obj(T.reshape(-1), EX + EdZ + EiZ)).max(), 0.0)
        nXZ = float(np.abs(Xb[blk]).max()) * float(np.abs(exact.ldexp_obj(dZm.reshape(-1), EdZ)).max()) * \
            float(np.abs(iZb[blk]).max())
        scale = (bm * float(np.abs(iZb[blk]).max()) + float(np.abs(Xb[blk]).max()) + n * nXZ + nT)
        two = Fraction(2)
        for j in range(n):
            for i_ in range(n):
                def raw(p, q):
                    return (bmF * Fraction(int(iZm[p, q])) * two ** EiZ - Fraction(int(Xm[p, q])) * two ** EX
                            - Fraction(int(T[p, q])) * two ** (EX + EdZ + EiZ))
                ex = (raw(i_, j) + raw(j, i_)) / 2
                err = max(err, float(abs(dXd[o + i_ + n * j] - ex)) / scale)
    chk.le("dX(dZ)", err, 2.0 ** -(P - 16) * nmax, "normwise relative to the terms")

    # forward errors against the 512-bit oracle (conditioning-dependent)
    for key in ("L", "dy", "dZ", "dX"):
        got = out[key]
        if key == "L":
            got = got[fx["ref_L_idx"] if "ref_L_idx" in fx else _lower_positions(m)]
        e = normwise(sdpg.unpack(got, L), sdpg.unpack(fx["ref_" + key], Lr))
        e256 = float(fx["err256_" + key][0])
        chk.le(key + " fwd", e, forward_bar(P, e256), f"oracle-256 {e256:.3e}")

    assert not chk.fail, f"{name} gemm={gemm}: " + "; ".join(chk.fail)
