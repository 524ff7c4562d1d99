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
