"""The Ozaki slice-count autotuner (sdpg_tune_slices) and the runs it serves.

The reference runs every product at its full mpf width (its precision is a run parameter,
sdpa_parts.cpp:199-200); the Ozaki form needs only as many slices as the request: 8S - 2 fraction
bits per row / column.  The tuner measures the effective bits of each built slice count against an
L + 2 limb product and its time, and picks the fastest count with at least request + 8 bits.  Checked:
  - the measured effective bits grow with S and are at least 8S - 12 (the GEMM bar of the Ozaki tests);
  - the choices for 200-, 256- and 320-bit requests are the smallest counts with 8S - 12 >= request + 8
    (28, 36, 44: a 200-bit request runs on 320-bit limbs as GMP rounds it, but needs 28 slices, not 36);
  - the dense rank-2 workload (n = 150, m = 1326) at 320 bits: both GEMM forms reach pdOPT at relative
    gap <= 1e-60 with iteration counts within 1 and objectives within 1e-50 of each other, the Ozaki form
    with the tuned count.
"""
import os
from decimal import Decimal

import pytest

import sdpg

PARAM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "sdpa-gmp_amd", "data",
                     "param_256_gap1e-60.sdpa")


@pytest.mark.gpu
@pytest.mark.parametrize("prec,want", [(200, 28), (256, 36), (320, 44)])
def test_tuner_picks_smallest_sufficient_slice_count(prec, want):
    print(f"case tune prec={prec}", flush=True)
    S, rows = sdpg.tune_slices(prec)
    for r in rows:
        print(f"  S={int(r[0])}: {r[1]:.1f} effective bits, {r[2]:.3f} ms (512^3), {r[3]:.3f} ms (1024^2 x 32)")
    assert S == want, (S, rows)
    bits = [r[1] for r in rows]
    assert all(b >= 8 * r[0] - 12 for b, r in zip(bits, rows)), rows
    assert bits == sorted(bits), rows


@pytest.mark.gpu
def test_dense_rank2_320bit_forms_agree(workload):
    runs = {}
    for name, gemm in (("plain", sdpg.GEMM_PLAIN), ("ozaki", sdpg.GEMM_OZAKI)):
        print(f"case w3 at 320 bits gemm={name}", flush=True)
        r = sdpg.solve(workload(3), PARAM, 320, None, gemm)
        print(f"  {r.iterations} iterations, {r.phase.decode().strip()}, relgap {r.relgap.decode()}, "
              f"slices {r.slices}, schur+chol {(r.t_schur_ms + r.t_chol_ms) / max(r.iterations, 1):.2f} ms/iter",
              flush=True)
        assert r.phase.decode().strip() == "pdOPT"
        assert Decimal(r.relgap.decode()) <= Decimal("1e-60")
        runs[name] = r
    p, o = runs["plain"], runs["ozaki"]
    assert o.slices == 44 and p.slices == 0
    assert abs(p.iterations - o.iterations) <= 1
    rel = lambda a, b: abs(Decimal(a) - Decimal(b)) / max(abs(Decimal(b)), Decimal(1))  # noqa: E731
    assert rel(o.objP70.decode(), p.objP70.decode()) < Decimal("1e-50")
    assert rel(o.objD70.decode(), p.objD70.decode()) < Decimal("1e-50")
