"""The batched Newton step (sdpg_batch_*) on problems of the batch workload (1-3 dense blocks of size
U[8,40], m U[10,120]) against each context's own engine on the same point.

Per problem, on a random interior point (X, Z symmetric positive definite, inv(Z) a symmetric positive
definite approximation of Z's inverse, y random): the engine's own predictor (sdpg_schur, sdpg_potrf,
sdpg_potrs) gives B_e and dy_e, its corrector (sdpg_gvec with corrector set, sdpg_potrs) dy_c.  The
batch then runs the same steps for all problems in lockstep, and its results are checked in exact
integer arithmetic with bars set by the products' precision P = 32L (element form):
  - B (strictly upper triangle, which the batched factorisation leaves in place) against B_e entry by
    entry, relative to sum_l ||A_i^l||_1 ||A_j^l||_1 ||X^l||_max ||inv(Z)^l||_max;
  - the factor (lower triangle): D (L L^T - B_e) D with D = diag(B_e)^(-1/2) rounded to powers of two,
    applied to two random sign vectors;
  - the predictor and corrector solves: D (B_e dy - g), relative to m |D^-1 dy|;
  - dy against dy_e normwise within 2^-200 (the step's forward error at these points is far smaller);
  - a problem requested with SDPG_BATCH_NONE keeps its B and dy bit for bit; a problem whose Schur
    matrix is not positive definite (X replaced by -X) reports info = 1, as the engine's sdpg_potrf does,
    and the others are unaffected;
  - graph replays: with sdpg_batch_set_graph on, three predictor replays with the contexts' own eager
    work in between give B and dy bit for bit equal to the eager batched call.
"""
import math
import os
import tempfile
from fractions import Fraction

import numpy as np
import pytest

import exact
import gen_problems
import sdpg

NPROB = 20
SKIP, FAIL = 3, 5
PREC = 256


def _point(rng, sizes):
    X, Z, iZ = [], [], []
    for n in sizes:
        for dst in (X, Z):
            q, _ = np.linalg.qr(rng.standard_normal((n, n)))
            a = (q * rng.uniform(0.5, 2.0, n)) @ q.T
            dst.append(np.triu(a) + np.triu(a, 1).T)
        b = np.linalg.inv(Z[-1])
        iZ.append(np.triu(b) + np.triu(b, 1).T)
    flat = lambda ms: [Fraction(float(v)) for mtx in ms for v in mtx.reshape(-1, order="F")]  # noqa: E731
    return flat(X), flat(Z), flat(iZ)


class Prob:
    def __init__(self, path, rng):
        self.path = path
        self.ctx = sdpg.Context(path, PREC)
        self.L = L = self.ctx.L
        self.m, self.sizes, n_lp, _, _, Am = exact.read_problem(path)
        assert n_lp == 0
        X, Z, iZ = _point(rng, self.sizes)
        y = [Fraction(float(v)) for v in rng.standard_normal(self.m)]
        self.X, self.Z, self.iZ, self.y = (sdpg.pack(v, L) for v in (X, Z, iZ, y))
        self.negX = sdpg.pack([-v for v in X], L)
        self.mu = sdpg.pack([Fraction(1, 4)], L)[0]
        self.beta = sdpg.pack([Fraction(1, 10)], L)[0]
        nb = len(self.sizes)
        self.A1 = np.zeros((self.m, nb))
        for i, ents in enumerate(Am):
            for blk, r, c, v in ents:
                self.A1[i, blk] += abs(float(v)) * (1 if r == c else 2)
        off = np.cumsum([0] + [n * n for n in self.sizes])
        xf, izf = exact.to_float(self.X, L), exact.to_float(self.iZ, L)
        self.nX = [float(np.abs(xf[off[b]:off[b + 1]]).max()) for b in range(nb)]
        self.niZ = [float(np.abs(izf[off[b]:off[b + 1]]).max()) for b in range(nb)]

    def prepare(self, corrector=False, negate=False):
        c = self.ctx
        if not corrector:
            c.set_point(self.negX if negate else self.X, self.Z, self.iZ, self.y)
            c.residuals()
        c.gvec(corrector, self.mu, self.beta, sdpg.SDPG_NOINFO)

    def engine_reference(self):
        c = self.ctx
        self.prepare()
        c.schur()
        self.Be = c.get(sdpg.SDPG_B)
        assert c.potrf() == 0
        c.potrs()
        self.dy_e = c.get(sdpg.SDPG_DY)
        c.dz(sdpg.SDPG_NOINFO)
        c.dx()
        self.prepare(corrector=True)
        c.potrs()
        self.dyc_e = c.get(sdpg.SDPG_DY)


def _pow2_scale(bdiag):
    return np.array([int(math.floor(math.log2(v) / 2)) if v > 0 else 0 for v in bdiag], dtype=np.int64)


def _scaled(vec_ints, E, k):
    return np.array([float(Fraction(int(v)) * Fraction(2) ** (E - int(ki))) if v else 0.0
                     for v, ki in zip(vec_ints, k)])


def _check_factor_and_solve(pr, Bb, dy, g, fails, tag):
    L, m = pr.L, pr.m
    P = 32 * L
    Bi, EB = exact.to_ints(pr.Be, L)
    Bmat = Bi.reshape(m, m, order="F")
    Bf = exact.ldexp_obj(Bi, EB).reshape(m, m, order="F")
    kD = _pow2_scale(np.diag(Bf))
    if tag == "predictor":
        # B above the diagonal against the engine's
        bi, Eb = exact.to_ints(Bb, L)
        up = np.array([i + m * j for j in range(m) for i in range(j)], dtype=np.int64)
        if len(up):
            d = exact.diff_float(bi[up], Eb, Bi[up], EB)
            ii, jj = up % m, up // m
            scale = np.array([sum(pr.A1[i, b] * pr.A1[j, b] * pr.nX[b] * pr.niZ[b] for b in range(len(pr.sizes)))
                              for i, j in zip(ii, jj)])
            err = float(np.max(np.abs(d) / np.maximum(scale, 1e-300)))
            bar = 2.0 ** -(P - 17) * max(pr.sizes) ** 2
            if err > bar:
                fails.append(f"{pr.path}: B {err:.3e} > {bar:.3e}")
        # factor: D (L L^T - B_e) D on two sign vectors
        lo = np.tril(np.ones((m, m), dtype=bool))
        Lmat = np.where(lo, bi.reshape(m, m, order="F"), 0).astype(object)
        rng = np.random.default_rng(11)
        kmax = int(kD.max())
        err = 0.0
        for _ in range(2):
            w = rng.choice([-1, 1], size=m)
            v = np.array([int(w[i]) << (kmax - int(kD[i])) for i in range(m)], dtype=object)
            r1 = Lmat.dot(Lmat.T.dot(v))
            r2 = Bmat.dot(v)
            a, b, E = exact.align(r1, 2 * Eb - kmax, r2, EB - kmax)
            err = max(err, float(np.max(np.abs(_scaled(a - b, E, kD)))))
        bar = 2.0 ** -(P - 16) * m * m
        if err > bar:
            fails.append(f"{pr.path}: LL^T-B {err:.3e} > {bar:.3e}")
    # solve: D (B_e dy - g)
    dyi, Edy = exact.to_ints(dy, L)
    gi, Eg = exact.to_ints(g, L)
    a, b, E = exact.align(Bmat.dot(dyi), EB + Edy, gi, Eg)
    rs = float(np.max(np.abs(_scaled(a - b, E, kD))))
    dyf = exact.ldexp_obj(dyi, Edy)
    bar = 2.0 ** -(P - 16) * m * m * float(np.max(np.abs(dyf * np.ldexp(1.0, kD))))
    if rs > bar:
        fails.append(f"{pr.path}: {tag} B dy - g {rs:.3e} > {bar:.3e}")


def _normwise(a, b, L):
    x, y = sdpg.unpack(a, L), sdpg.unpack(b, L)
    s = max(abs(v) for v in y)
    return float(max(abs(p - q) for p, q in zip(x, y)) / s) if s else float(max(abs(v) for v in x))


@pytest.fixture(scope="module")
def probs():
    tmp = tempfile.mkdtemp(prefix="sdpg_batch_t_")
    # the first NPROB problems whose m constraint matrices can be independent (m <= sum n(n+1)/2; in the
    # others B is singular at every point, as for the reference)
    paths = []
    for p in gen_problems.generate_batch(tmp, 2 * NPROB):
        m, sizes = exact.read_problem(p)[:2]
        if m <= sum(n * (n + 1) // 2 for n in sizes):
            paths.append(p)
    paths = paths[:NPROB]
    assert len(paths) == NPROB
    rng = np.random.default_rng(20260)
    out = [Prob(p, rng) for p in paths]
    for pr in out:
        print(f"engine reference {os.path.basename(pr.path)} m={pr.m} blocks={pr.sizes}", flush=True)
        pr.engine_reference()
    yield out
    for pr in out:
        pr.ctx.close()


@pytest.mark.gpu
def test_batch_matches_engine(probs):
    batch = sdpg.Batch([p.ctx for p in probs])
    try:
        # predictor for all but SKIP; FAIL gets -X
        for k, pr in enumerate(probs):
            pr.prepare(negate=(k == FAIL))
        keep_B, keep_dy = probs[SKIP].ctx.get(sdpg.SDPG_B), probs[SKIP].ctx.get(sdpg.SDPG_DY)
        req = np.full(len(probs), sdpg.BATCH_PREDICTOR, dtype=np.int32)
        req[SKIP] = sdpg.BATCH_NONE
        print("batched predictor", flush=True)
        info = batch.newton(req)
        print(f"info {info.tolist()}, {batch.time_ms():.3f} ms", flush=True)
        assert info[FAIL] == 1, f"non-positive-definite Schur matrix: info {info[FAIL]}"
        assert all(info[k] == 0 for k in range(len(probs)) if k != FAIL)
        assert np.array_equal(probs[SKIP].ctx.get(sdpg.SDPG_B), keep_B)
        assert np.array_equal(probs[SKIP].ctx.get(sdpg.SDPG_DY), keep_dy)
        fails = []
        ok = [k for k in range(len(probs)) if k not in (SKIP, FAIL)]
        for k in ok:
            pr = probs[k]
            c = pr.ctx
            dy, g = c.get(sdpg.SDPG_DY), c.get(sdpg.SDPG_GVEC)
            _check_factor_and_solve(pr, c.get(sdpg.SDPG_B), dy, g, fails, "predictor")
            fw = _normwise(dy, pr.dy_e, pr.L)
            if fw > 2.0 ** -200:
                fails.append(f"{pr.path}: dy forward {fw:.3e}")
            c.dz(sdpg.SDPG_NOINFO)
            c.dx()
            pr.prepare(corrector=True)
        req = np.full(len(probs), sdpg.BATCH_NONE, dtype=np.int32)
        req[ok] = sdpg.BATCH_CORRECTOR
        print("batched corrector", flush=True)
        info = batch.newton(req)
        assert not info.any()
        for k in ok:
            pr = probs[k]
            c = pr.ctx
            dy, g = c.get(sdpg.SDPG_DY), c.get(sdpg.SDPG_GVEC)
            _check_factor_and_solve(pr, None, dy, g, fails, "corrector")
            fw = _normwise(dy, pr.dyc_e, pr.L)
            if fw > 2.0 ** -200:
                fails.append(f"{pr.path}: corrector dy forward {fw:.3e}")
        # the engine reports the same failure on the -X point
        probs[FAIL].prepare(negate=True)
        probs[FAIL].ctx.schur()
        assert probs[FAIL].ctx.potrf() == 1
        assert not fails, "\n".join(fails)
    finally:
        batch.close()


@pytest.mark.gpu
def test_batch_graph_replays(probs):
    batch = sdpg.Batch([p.ctx for p in probs])
    try:
        req = np.full(len(probs), sdpg.BATCH_PREDICTOR, dtype=np.int32)
        for pr in probs:
            pr.prepare()
        assert not batch.newton(req).any()
        ref = [(pr.ctx.get(sdpg.SDPG_B), pr.ctx.get(sdpg.SDPG_DY)) for pr in probs]
        eager_ms = batch.time_ms()
        batch.set_graph(True)
        for rep in range(3):
            for pr in probs:
                pr.prepare()
            assert not batch.newton(req).any()
            print(f"replay {rep}: {batch.time_ms():.3f} ms (eager {eager_ms:.3f} ms)", flush=True)
            for pr, (B, dy) in zip(probs, ref):
                assert np.array_equal(pr.ctx.get(sdpg.SDPG_B), B), f"replay {rep}: B differs for {pr.path}"
                assert np.array_equal(pr.ctx.get(sdpg.SDPG_DY), dy), f"replay {rep}: dy differs for {pr.path}"
    finally:
        batch.close()


def _rel_matrix(a, b, L):
    """normwise relative difference of two packed arrays (exact, as a float)"""
    ai, Ea = exact.to_ints(a, L)
    bi, Eb = exact.to_ints(b, L)
    d = exact.diff_float(ai, Ea, bi, Eb)
    s = float(np.max(np.abs(exact.to_float(b, L))))
    return float(np.max(np.abs(d))) / s if s else float(np.max(np.abs(d)))


@pytest.mark.gpu
def test_batch_phases_match_contexts(probs):
    """Every batched phase of an iteration against the context's own call on the same point: the inverse
    factors, residuals, Make_gVec's g (predictor in an infeasible and a feasible phase), dZ and dX from
    the same dy, the ten inner products (against their exact values from the device's arrays), the
    step-length eigenvalues and the update."""
    L = probs[0].L
    P = 32 * L
    n = len(probs)
    phases = np.array([sdpg.SDPG_NOINFO if k % 2 == 0 else sdpg.SDPG_PDFEAS for k in range(n)], dtype=np.int32)
    aP = [Fraction(7, 10), Fraction(1, 3)]
    aD = [Fraction(9, 10), Fraction(1, 5)]
    ref = []
    for k, pr in enumerate(probs):
        c = pr.ctx
        c.set_point(pr.X, pr.Z, pr.iZ, pr.y)
        assert c.factor_point() == 0
        r = {"icX": c.get(sdpg.SDPG_INVCHOLX), "icZ": c.get(sdpg.SDPG_INVCHOLZ), "invZ": c.get(sdpg.SDPG_INVZ)}
        r["res"] = c.residuals()
        c.gvec(False, pr.mu, pr.beta, int(phases[k]))
        r["g"] = c.get(sdpg.SDPG_GVEC)
        c.schur()
        assert c.potrf() == 0
        c.potrs()
        r["dy"] = c.get(sdpg.SDPG_DY)
        c.dz(int(phases[k]))
        r["dZ"] = c.get(sdpg.SDPG_DZ)
        c.dx()
        r["dX"] = c.get(sdpg.SDPG_DX)
        r["ip"] = [c.inner(w) for w in range(1, 11)]
        r["eig"] = [c.min_eigen(0), c.min_eigen(1)]
        c.update(sdpg.pack([aP[k % 2]], L)[0], sdpg.pack([aD[k % 2]], L)[0])
        r["X1"], r["Z1"], r["y1"] = c.get(sdpg.SDPG_X), c.get(sdpg.SDPG_Z), c.get(sdpg.SDPG_Y)
        ref.append(r)
    batch = sdpg.Batch([p.ctx for p in probs])
    fails = []

    def close(tag, k, got, want, bar):
        err = _rel_matrix(got, want, L)
        if err > bar:
            fails.append(f"{os.path.basename(probs[k].path)} {tag}: {err:.3e} > {bar:.3e}")

    try:
        on = np.ones(n, dtype=np.int32)
        for pr in probs:
            pr.ctx.set_point(pr.X, pr.Z, pr.iZ, pr.y)
        print("batched factor_point / residuals / gvec", flush=True)
        assert not batch.factor_point(on).any()
        for k, pr in enumerate(probs):
            for key, sel in (("icX", sdpg.SDPG_INVCHOLX), ("icZ", sdpg.SDPG_INVCHOLZ), ("invZ", sdpg.SDPG_INVZ)):
                close(key, k, pr.ctx.get(sel), ref[k][key], 2.0 ** -(P - 40))
        res = batch.residuals(on)
        for k in range(n):
            for q in range(4):
                close(f"residual {q}", k, res[k, q], ref[k]["res"][q], 2.0 ** -(P - 40))
        mus = np.stack([pr.mu for pr in probs])
        betas = np.stack([pr.beta for pr in probs])
        batch.gvec(on, mus, betas, phases)
        for k, pr in enumerate(probs):
            close("g", k, pr.ctx.get(sdpg.SDPG_GVEC), ref[k]["g"], 2.0 ** -(P - 40))
        assert not batch.newton(on).any()
        for k, pr in enumerate(probs):
            pr.ctx.put(sdpg.SDPG_DY, ref[k]["dy"])  # the same dy for dZ and dX
        print("batched dz / dx / inner / min_eigen / update", flush=True)
        batch.dz(on, phases)
        for k, pr in enumerate(probs):
            close("dZ", k, pr.ctx.get(sdpg.SDPG_DZ), ref[k]["dZ"], 2.0 ** -(P - 40))
        batch.dx(on)
        for k, pr in enumerate(probs):
            close("dX", k, pr.ctx.get(sdpg.SDPG_DX), ref[k]["dX"], 2.0 ** -(P - 40))
        ips = batch.inner(on)
        for k, pr in enumerate(probs):
            c = pr.ctx
            arr = {w: exact.to_ints(c.get(w), L) for w in (sdpg.SDPG_X, sdpg.SDPG_Z, sdpg.SDPG_DX, sdpg.SDPG_DZ,
                                                           sdpg.SDPG_Y, sdpg.SDPG_DY)}
            Cm = exact.read_problem(pr.path)[4]
            Cd = [np.full((nb, nb), Fraction(0), dtype=object) for nb in pr.sizes]
            for blk, i, j, v in Cm:
                Cd[blk][i, j] = v
                Cd[blk][j, i] = v
            Cflat = [v for mtx in Cd for v in mtx.reshape(-1, order="F")]
            bvec = exact.read_problem(pr.path)[3]

            def val(w):
                v, E = arr[w]
                return [Fraction(int(x)) * Fraction(2) ** E for x in v]

            X, Z, dX, dZ, y, dy = (val(w) for w in (sdpg.SDPG_X, sdpg.SDPG_Z, sdpg.SDPG_DX, sdpg.SDPG_DZ,
                                                    sdpg.SDPG_Y, sdpg.SDPG_DY))
            offs = np.cumsum([0] + [nb * nb for nb in pr.sizes])
            diag = [int(offs[b]) + i + nb * i for b, nb in enumerate(pr.sizes) for i in range(nb)]
            pairs = {1: (X, Z), 2: (X, dZ), 3: (dX, Z), 4: (dX, dZ), 5: (Cflat, X), 6: (Cflat, dX), 7: (bvec, y),
                     8: (bvec, dy)}
            for w in range(1, 11):
                if w in pairs:
                    a, b = pairs[w]
                    exact_v = sum(p_ * q_ for p_, q_ in zip(a, b))
                    scale = float(sum(abs(p_ * q_) for p_, q_ in zip(a, b)))
                else:
                    src = X if w == sdpg.SDPG_IP_TRX else Z
                    exact_v = sum(src[i] for i in diag)
                    scale = float(sum(abs(src[i]) for i in diag))
                got = sdpg.unpack(ips[k, w - 1], L)[0]
                err = float(abs(got - exact_v))
                if err > 2.0 ** -(P - 16) * max(scale, 1e-300):
                    fails.append(f"{os.path.basename(pr.path)} inner {w}: {err:.3e} vs scale {scale:.3e}")
        for which in (0, 1):
            eg = batch.min_eigen(on, which)
            for k in range(n):
                got = float(sdpg.unpack(eg[k], L)[0])
                want = float(sdpg.unpack(ref[k]["eig"][which], L)[0])
                if abs(got - want) > 1e-10 * max(abs(want), 1.0):
                    fails.append(f"{os.path.basename(probs[k].path)} min_eigen {which}: {got!r} vs {want!r}")
        for k, pr in enumerate(probs):
            pr.ctx.put(sdpg.SDPG_DX, ref[k]["dX"])
            pr.ctx.put(sdpg.SDPG_DZ, ref[k]["dZ"])
        batch.update(on, sdpg.pack([aP[k % 2] for k in range(n)], L), sdpg.pack([aD[k % 2] for k in range(n)], L))
        for k, pr in enumerate(probs):
            for key, sel in (("X1", sdpg.SDPG_X), ("Z1", sdpg.SDPG_Z), ("y1", sdpg.SDPG_Y)):
                got = pr.ctx.get(sel)
                if not np.array_equal(got, ref[k][key]):
                    close(key, k, got, ref[k][key], 2.0 ** -(P - 8))
        print(f"batched phase timers (ms): {[round(t, 3) for t in batch.timers()]}", flush=True)
        assert not fails, "\n".join(fails)
    finally:
        batch.close()
