"""Newton-step parity on the GPU: the engine's Rp, Rd, g, B, L, dy, dX, dZ on the oracle's
iterates against the 512-bit oracle on the same iterate.

Bar: normwise error (max |got - ref| / max |ref|) at most 64x the 256-bit oracle's own error
on that iterate, floored at 2^-240 (the 256-bit arithmetic itself).  Residuals (cancellation
quantities) are compared in absolute terms against the size of the iterate.
"""
import glob
import os
from fractions import Fraction

import numpy as np
import pytest

import sdpg

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(GOLD, "newton_w*_it*.npz")))
FLOOR = 2.0 ** -240


def _fr(arr, L):
    return sdpg.unpack(arr, L)


def normwise(got, ref):
    scale = max(abs(x) for x in ref)
    if scale == 0:
        return float(max(abs(x) for x in got))
    return float(max(abs(a - b) for a, b in zip(got, ref)) / scale)


def run_step(ctx, fx):
    L = ctx.L
    rp = lambda k: sdpg.repack(fx["in_" + k], fx["in_" + k].shape[1] - 2, L)  # noqa: E731
    ctx.set_point(rp("X"), rp("Z"), rp("invZ"), rp("y"))
    ctx.residuals()
    phase = int(fx["phase"][0])
    ctx.gvec(False, rp("mu")[0], rp("beta")[0], phase)
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


@pytest.mark.gpu
@pytest.mark.parametrize("gemm", [sdpg.GEMM_PLAIN, sdpg.GEMM_OZAKI], ids=["plain", "ozaki"])
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[7:-4] for p in FIXTURES])
def test_newton_step_matches_oracle(path, gemm, workload):
    name = os.path.basename(path)
    wk = int(name.split("_")[1][1:])
    fx = dict(np.load(path))
    print(f"case {name} gemm={gemm}", flush=True)
    ctx = sdpg.Context(workload(wk), 256, gemm=gemm)
    try:
        out = run_step(ctx, fx)
    finally:
        ctx.close()
    L = ctx.L
    Lref = fx["ref_dy"].shape[1] - 2
    X = _fr(fx["in_X"], fx["in_X"].shape[1] - 2)
    Z = _fr(fx["in_Z"], fx["in_Z"].shape[1] - 2)
    size = max(abs(v) for v in X + Z)
    for key in ("Rp", "Rd"):
        got, ref = _fr(out[key], L), _fr(fx["ref_" + key], Lref)
        err = max(abs(a - b) for a, b in zip(got, ref))
        assert err <= Fraction(1, 10 ** 60) * (1 + size), f"{key}: absolute error {float(err):.3e}"
    for key in ("g", "B", "L", "dy", "dX", "dZ"):
        if "ref_" + key not in fx:
            continue
        got = out[key]
        if key == "L":
            m = int(round(len(got) ** 0.5))
            mask = np.array([i >= j for j in range(m) for i in range(m)])
            got = got[mask]
        e = normwise(_fr(got, L), _fr(fx["ref_" + key], Lref))
        tol = max(64.0 * float(fx["err256_" + key][0]), FLOOR)
        print(f"  {key}: normwise error {e:.3e} (oracle-256 {float(fx['err256_' + key][0]):.3e}, tol {tol:.3e})")
        assert e <= tol, f"{key}: normwise error {e:.3e} > {tol:.3e}"
