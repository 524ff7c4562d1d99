"""End-to-end batched form: lockstep solves (sdpa_amd_batch_solve) of the first problems of the seeded
batch workload (1-3 dense blocks of U[8,40], m U[10,120], 256 bits, epsilonStar = epsilonDash = 1e-60).

  - against independent solves of the same problems (sdpa_amd_solve, one context each): the same phase,
    iteration counts within 1, objectives to 1e-50 relative and relative gap <= 1e-60 where pdOPT;
  - against the 512-bit oracle on the problems of tests/golden/batch_solves.json (blocks on both sides of
    the eigenvalue routines' n = 20 switch): pdOPT, iteration count within 1, objectives to 1e-50;
  - problems whose m exceeds the dimension of their symmetric space (B singular at every point) stop at
    iteration 0 with the Cholesky failure, in both forms.
"""
import concurrent.futures
import hashlib
import json
import os
import tempfile
from decimal import Decimal

import pytest

import exact
import gen_problems
import sdpg

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PARAM = os.path.join(os.path.dirname(GOLD), "..", "sdpa-gmp_amd", "data", "param_256_gap1e-60.sdpa")
NPROB = 24


def rel(a, b):
    a, b = Decimal(a), Decimal(b)
    return abs(a - b) / max(abs(b), Decimal(1))


@pytest.mark.gpu
def test_lockstep_solves_match_independent_and_oracle():
    tmp = tempfile.mkdtemp(prefix="sdpg_bs_")
    paths = gen_problems.generate_batch(tmp, NPROB)
    print(f"lockstep solve of {NPROB} problems", flush=True)
    res, lat, stats = sdpg.batch_solve(paths, PARAM, 0, 0, True)
    print(f"  {stats}", flush=True)

    def one(p):
        return sdpg.solve(p, PARAM, 0, None, sdpg.GEMM_PLAIN)

    print("independent solves", flush=True)
    with concurrent.futures.ThreadPoolExecutor(8) as ex:
        ind = list(ex.map(one, paths))
    with open(os.path.join(GOLD, "batch_solves.json")) as f:
        gold = json.load(f)
    fails = []
    for k, (a, b) in enumerate(zip(res, ind)):
        pa, pb = a.phase.decode().strip(), b.phase.decode().strip()
        m, sizes = exact.read_problem(paths[k])[:2]
        tag = f"problem {k} (m={m}, blocks {sizes})"
        print(f"  {tag}: lockstep {pa} {a.iterations}, independent {pb} {b.iterations}, latency {lat[k]:.2f} s",
              flush=True)
        if m > sum(n * (n + 1) // 2 for n in sizes):
            if not (a.iterations == 0 and b.iterations == 0 and a.status == 1 and b.status == 1):
                fails.append(f"{tag}: singular Schur matrix should stop at iteration 0")
            continue
        if pa != pb or abs(a.iterations - b.iterations) > 1:
            fails.append(f"{tag}: {pa}/{a.iterations} vs {pb}/{b.iterations}")
            continue
        if pa == "pdOPT":
            if Decimal(a.relgap.decode()) > Decimal("1e-60"):
                fails.append(f"{tag}: relgap {a.relgap.decode()}")
            for key in ("objP70", "objD70"):
                d = rel(getattr(a, key).decode(), getattr(b, key).decode())
                if d > Decimal("1e-50"):
                    fails.append(f"{tag}: {key} differs by {d:.2e}")
        g = gold.get(str(k))
        if g:
            with open(paths[k], "rb") as f:
                assert hashlib.sha256(f.read()).hexdigest() == g["sha256"], f"problem {k} input drifted"
            o = g["oracle512"]
            if pa != o["phase"] or abs(a.iterations - o["iterations"]) > 1:
                fails.append(f"{tag}: {pa}/{a.iterations} vs oracle {o['phase']}/{o['iterations']}")
            for key, ok in (("objP70", "objP"), ("objD70", "objD")):
                d = rel(getattr(a, key).decode(), o[ok])
                if d > Decimal("1e-50"):
                    fails.append(f"{tag}: {key} differs from the oracle by {d:.2e}")
    assert all(str(k) in gold for k in (2, 11, 14, 17, 20))
    assert not fails, "\n".join(fails)
