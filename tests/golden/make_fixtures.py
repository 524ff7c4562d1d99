"""Regenerate the committed golden fixtures under tests/golden/.

Run in the development container (needs /root/reference for the SDPLIB goldens):
    python tests/golden/make_fixtures.py [--skip-sdplib] [--skip-newton]

Produces
  inputs.json             sha256 of the generated workload files (gen_problems.py)
  newton_w{0,1}_it*.npz   Newton-step fixtures: the oracle's 256-bit iterate (inputs, exact
                          in 12 interchange limbs) and the 512-bit oracle outputs on that same
                          iterate (16 limbs), with the 256-bit oracle's normwise error per output;
                          workload 1 keeps SAMPLE entries of B and of L's lower triangle (indices in
                          ref_B_idx / ref_L_idx, column-major positions in the m x m matrix)
  solves.json             end-to-end oracle summaries of workloads 0 and 1 at epsilonStar =
                          epsilonDash = 1e-60: at 256 bits (GMP truncation: the primal residual
                          stalls near 1e-49, so the run ends pFEAS) and at 512 bits (pdOPT; the
                          reference trajectory for the GPU's 256-bit run, objectives to 70 digits)
  batch_solves.json       512-bit oracle solves (epsilonStar = epsilonDash = 1e-60) of problems
                          BATCH_PICK of the seeded batch workload (gen_problems.generate_batch), with
                          the sha256 of each generated input
  sdplib/*.dat-s.xz, sdplib_golden.json
                          inputs and recorded results of the reference's own regression runs
                          (SDPLIB/data/*.result.xz, example1.result): iteration count, phase,
                          objectives, relative gap and the per-iteration log columns
"""
from __future__ import annotations

import argparse
import json
import lzma
import os
import re
import shutil
import sys
import tempfile
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "sdpa-gmp_amd"))

import gen_problems  # noqa: E402
import oracle_lib  # noqa: E402

PARAM = os.path.join(ROOT, "sdpa-gmp_amd", "data", "param_256_gap1e-60.sdpa")
REF = "/root/reference"

SAMPLE = 3000  # sampled entries of B and L kept for workload 1 (m = 351: the full matrices are ~9 MB each)

SDPLIB = ["control1", "control2", "truss1", "truss4", "hinf1", "hinf3", "theta1", "infp1", "infd1", "qap5"]


def words_to_frac(w):
    L = len(w) - 2
    M = 0
    for k in range(L - 1, -1, -1):
        M = (M << 32) | int(w[2 + k])
    if M == 0:
        return Fraction(0)
    e = int(np.int32(np.uint32(w[0])))
    sh = e - 32 * L
    v = Fraction(M << sh) if sh >= 0 else Fraction(M, 1 << -sh)
    return -v if int(w[1]) & 1 else v


def normwise_err(a, ref):
    fa = [words_to_frac(r) for r in a]
    fr = [words_to_frac(r) for r in ref]
    scale = max((abs(x) for x in fr), default=Fraction(0))
    if scale == 0:
        return 0.0
    err = max(abs(x - y) for x, y in zip(fa, fr))
    return float(err / scale)


def gen_inputs(tmp):
    out = {}
    for k in (0, 1):
        path = os.path.join(tmp, f"w{k}.dat-s")
        out[f"w{k}"] = gen_problems.generate(k, path)
    return out


def newton_fixtures(tmp, name, iters):
    data = os.path.join(tmp, f"{name}.dat-s")
    dump = os.path.join(tmp, f"dump_{name}")
    res = oracle_lib.solve(data, PARAM, 0, os.path.join(tmp, f"{name}.log"), dump, iters, 12)
    summary = {"iterations": res.iterations, "phase": res.phase_name.decode().strip(),
               "objP": res.objP.decode(), "objD": res.objD.decode(), "relgap": res.relgap_s.decode(),
               "pfeas": res.pfeas_s.decode(), "dfeas": res.dfeas_s.decode()}
    with open(os.path.join(tmp, f"{name}.log")) as f:
        summary["log"] = [ln.rstrip("\n") for ln in f if re.match(r"^\s*\d+ ", ln)]
    for it in iters:
        d = os.path.join(dump, f"iter{it:03d}")
        if not os.path.isdir(d):
            continue
        o512 = os.path.join(tmp, f"{name}_o512_{it}")
        o256 = os.path.join(tmp, f"{name}_o256_{it}")
        rc5 = oracle_lib.newton_step(data, 512, d, o512, 16)
        rc2 = oracle_lib.newton_step(data, 256, d, o256, 16)
        if rc5 or rc2:
            print(f"skip {name} iter {it}: rc {rc5} {rc2}")
            continue
        arrs = {}
        for key in ("X", "Z", "invZ", "y", "mu", "beta"):
            arrs["in_" + key] = oracle_lib.read_dump(os.path.join(d, key + ".bin"))
        with open(os.path.join(d, "phase.txt")) as f:
            arrs["phase"] = np.array([int(f.read().strip())])
        for key in ("Rp", "Rd", "g", "dy", "dX", "dZ", "B", "L"):
            r512 = oracle_lib.read_dump(os.path.join(o512, key + ".bin"))
            r256 = oracle_lib.read_dump(os.path.join(o256, key + ".bin"))
            if key in ("B", "L"):
                m = int(round(len(r512) ** 0.5))
                if key == "L":  # only the lower triangle is the factor
                    pos = np.array([i + m * j for j in range(m) for i in range(j, m)], dtype=np.int64)
                else:
                    pos = np.arange(m * m, dtype=np.int64)
                if name != "w0":
                    rng = np.random.default_rng(1000 + it + (7 if key == "L" else 0))
                    diag = np.array([i + m * i for i in range(m)], dtype=np.int64)
                    pick = rng.choice(pos, size=min(SAMPLE, len(pos)), replace=False)
                    pos = np.unique(np.concatenate([pick, diag[:: max(1, m // 64)]]))
                    arrs["ref_" + key + "_idx"] = pos
                r512 = r512[pos]
                r256 = r256[pos]
            arrs["ref_" + key] = r512
            arrs["err256_" + key] = np.array([normwise_err(r256, r512)])
            print(f"{name} it{it} {key}: oracle-256 normwise error {arrs['err256_' + key][0]:.3e}", flush=True)
        np.savez_compressed(os.path.join(HERE, f"newton_{name}_it{it:03d}.npz"), **arrs)
    return summary


def solve_summary(tmp, name, prec):
    data = os.path.join(tmp, f"{name}.dat-s")
    log = os.path.join(tmp, f"{name}_{prec}.log")
    res = oracle_lib.solve(data, PARAM, prec, log)
    out = {"precision": prec, "iterations": res.iterations, "phase": res.phase_name.decode().strip(),
           "objP": res.objP70.decode(), "objD": res.objD70.decode(), "relgap": res.relgap_s.decode(),
           "pfeas": res.pfeas_s.decode(), "dfeas": res.dfeas_s.decode()}
    with open(log) as f:
        out["log"] = [ln.rstrip("\n") for ln in f if re.match(r"^\s*\d+ ", ln)]
    print(f"{name} at {prec} bits: {out['iterations']} iterations, {out['phase']}, relgap {out['relgap']}", flush=True)
    return out


BATCH_PICK = (2, 11, 14, 17, 20)  # blocks of 10..22 (both eigenvalue paths), m 10..31


def batch_solves():
    import hashlib

    import gen_problems

    tmp = tempfile.mkdtemp(prefix="sdpg_fix_b_")
    try:
        paths = gen_problems.generate_batch(tmp, max(BATCH_PICK) + 1)
        out = {}
        for k in BATCH_PICK:
            with open(paths[k], "rb") as f:
                sha = hashlib.sha256(f.read()).hexdigest()
            res = oracle_lib.solve(paths[k], PARAM, 512, None)
            out[str(k)] = {"sha256": sha, "oracle512": {
                "iterations": res.iterations, "phase": res.phase_name.decode().strip(), "objP": res.objP70.decode(),
                "objD": res.objD70.decode(), "relgap": res.relgap_s.decode()}}
            print(f"batch problem {k}: {res.iterations} iterations, {res.phase_name.decode().strip()}", flush=True)
        with open(os.path.join(HERE, "batch_solves.json"), "w") as f:
            json.dump(out, f, indent=1)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def parse_result(text):
    g = {}
    m = re.search(r"phase.value = (\S+)", text)
    g["phase"] = m.group(1)
    g["iterations"] = int(re.search(r"Iteration = (\d+)", text).group(1))
    for key, pat in (("relgap", r"relative gap = (\S+)"), ("objP", r"objValPrimal = (\S+)"),
                     ("objD", r"objValDual   = (\S+)"), ("mu", r"\n\s+mu = (\S+)")):
        mm = re.search(pat, text)
        g[key] = mm.group(1) if mm else None
    g["log"] = [ln for ln in text.splitlines() if re.match(r"^\s*\d+ \d\.\de[+-]\d\d ", ln)]
    return g


def sdplib_fixtures():
    out_dir = os.path.join(HERE, "sdplib")
    os.makedirs(out_dir, exist_ok=True)
    gold = {}
    for name in ("example1.dat-s", "example1.dat", "example1.ini", "example1.ini-s", "example2.dat",
                 "example1.result"):
        shutil.copy(os.path.join(REF, name), os.path.join(out_dir, name))
    with open(os.path.join(REF, "example1.result")) as f:
        gold["example1"] = parse_result(f.read())
    for name in SDPLIB:
        shutil.copy(os.path.join(REF, "SDPLIB", "data", f"{name}.dat-s.xz"), os.path.join(out_dir, f"{name}.dat-s.xz"))
        with lzma.open(os.path.join(REF, "SDPLIB", "data", f"{name}.result.xz"), "rt") as f:
            gold[name] = parse_result(f.read())
    with open(os.path.join(HERE, "sdplib_golden.json"), "w") as f:
        json.dump(gold, f, indent=1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-sdplib", action="store_true")
    ap.add_argument("--skip-newton", action="store_true")
    ap.add_argument("--solves-only", action="store_true", help="recompute solves.json only")
    ap.add_argument("--newton-only", choices=["w0", "w1"], help="recompute the Newton fixtures of one workload")
    ap.add_argument("--batch-solves", action="store_true", help="recompute batch_solves.json only")
    a = ap.parse_args()
    oracle_lib.build_oracle()
    if a.batch_solves:
        batch_solves()
        return
    if a.solves_only:
        tmp = tempfile.mkdtemp(prefix="sdpg_fix_")
        try:
            gen_inputs(tmp)
            solves = {w: {"oracle256": solve_summary(tmp, w, 256), "oracle512": solve_summary(tmp, w, 512)}
                      for w in ("w0", "w1")}
            with open(os.path.join(HERE, "solves.json"), "w") as f:
                json.dump(solves, f, indent=1)
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        return
    if a.newton_only:
        tmp = tempfile.mkdtemp(prefix="sdpg_fix_")
        try:
            gen_inputs(tmp)
            newton_fixtures(tmp, a.newton_only, {"w0": [0, 1, 5, 20, 40, 60], "w1": [0, 10, 45]}[a.newton_only])
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        return
    if not a.skip_sdplib:
        sdplib_fixtures()
    if a.skip_newton:
        return
    tmp = tempfile.mkdtemp(prefix="sdpg_fix_")
    try:
        hashes = gen_inputs(tmp)
        with open(os.path.join(HERE, "inputs.json"), "w") as f:
            json.dump(hashes, f, indent=1)
        newton_fixtures(tmp, "w0", [0, 1, 5, 20, 40, 60])
        newton_fixtures(tmp, "w1", [0, 10, 45])
        solves = {w: {"oracle256": solve_summary(tmp, w, 256), "oracle512": solve_summary(tmp, w, 512)}
                  for w in ("w0", "w1")}
        with open(os.path.join(HERE, "solves.json"), "w") as f:
            json.dump(solves, f, indent=1)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
