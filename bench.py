"""Benchmark: Schur complement build + Cholesky per IPM iteration at 256 bits (gap <= 1e-60),
Ozaki INT8-MFMA form against the plain limb form.

Workload (workload 2 of gen_problems.py): max-cut relaxation, n = m = 1000, one SDP block,
A_i = e_i e_i^T, b = 1, C = -Laplacian/4 of an Erdos-Renyi G(1000, 0.01) graph with weights
+-1; precision 256 bits (GMP limb rounding: 10 x 32-bit limbs), epsilonStar = epsilonDash = 1e-60.
The Schur matrix is dense 1000 x 1000 (F3 formula, B = X o inv(Z)).

For each GEMM form the solver first runs to convergence (untimed here; its per-iteration device
times over the whole trajectory are reported), then a second run stops at iteration `--at`
and a "step" = one Make_bMat + Cholesky (sdpg_schur + sdpg_potrf) on that iterate, resident in
HBM, is run `--warmup` times untimed and `--steps` times timed.  value = wall seconds per step of
the Ozaki form.

roofline: the Ozaki MFMA product kernel, INT8 operations = 2 * S(S+1)/2 per mp multiply-add of the
reference loop (S slices), over its event-timed average launch duration in the timed steps, against
the 5.0 POP/s dense INT8 peak.  The per-kernel breakdown of the step is reported beside it.

cpu_baseline: the oracle (C restatement of the reference's mpf path, serial Cholesky as the
reference's Rpotrf) on the same iterate, dumped by the GPU run: Make_bMat in full and the Cholesky of
the leading --cpu-n rows/columns of B, scaled by (m / cpu_n)^3 to the full factorisation.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "sdpa-gmp_amd")
sys.path.insert(0, PKG)
sys.path.insert(0, os.path.join(ROOT, "tests"))

PARAM = os.path.join(PKG, "data", "param_256_gap1e-60.sdpa")
INT8_PEAK = 5.0e15  # MI355X dense INT8 MFMA, ops/s
HBM_PEAK = 8.0e12


def structure(path: str) -> dict:
    """m, block sizes and the reference's mp multiply-add count of one Make_bMat + Cholesky."""
    with open(path) as f:
        lines = [ln for ln in f if not ln.startswith('"') and not ln.startswith("*")]
    m = int(lines[0].split()[0])
    nb = int(lines[1].split()[0])
    blocks = [int(x) for x in lines[2].replace("=", " ").split() if x.lstrip("-").isdigit()][:nb]
    return {"m": m, "blocks": blocks, "chol_mp_mac": m ** 3 / 6.0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--at", type=int, default=20, help="iteration whose iterate the timed steps replay")
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--forms", default="ozaki,plain")
    ap.add_argument("--no-solve", action="store_true", help="skip the full trajectories")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=500, help="Cholesky sample size of the CPU baseline")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (multi-GPU sharding of the "
                                   "Schur rows and a 2D block-cyclic Cholesky are design notes only)"}))
        sys.exit(2)

    import gen_problems
    import sdpg

    tmp = tempfile.mkdtemp(prefix="sdpg_bench_")
    data = os.path.join(tmp, f"w{a.workload}.dat-s")
    sha = gen_problems.generate(a.workload, data)
    info = structure(data)
    L = sdpg.limbs_for_bits(256)
    forms = {"ozaki": sdpg.GEMM_OZAKI, "plain": sdpg.GEMM_PLAIN}
    res = {}
    dump = os.path.join(tmp, "it_")
    for idx, name in enumerate(a.forms.split(",")):
        mode = forms[name]
        r_out = {}
        if not a.no_solve:
            t0 = time.time()
            r = sdpg.solve(data, PARAM, 0, None, mode)
            t_solve = time.time() - t0
            it = max(r.iterations, 1)
            r_out = {"iterations": r.iterations, "phase": r.phase.decode().strip(), "relgap": r.relgap.decode(),
                     "objP": r.objP.decode(), "objD": r.objD.decode(), "solve_wall_s": t_solve,
                     "traj_ms_per_iter": {"schur": r.t_schur_ms / it, "chol": r.t_chol_ms / it,
                                          "solve": r.t_solve_ms / it, "gvec": r.t_gvec_ms / it,
                                          "dxdz": r.t_dxdz_ms / it, "inverse": r.t_inv_ms / it,
                                          "min_eigen": r.t_eig_ms / it}}
            print(f"[bench] {name}: solve {r.iterations} iterations, {r_out['phase']}, relgap {r_out['relgap']}, "
                  f"{t_solve:.1f} s; schur+chol {(r.t_schur_ms + r.t_chol_ms) / it:.2f} ms/iter over the trajectory",
                  flush=True)
        ms_s, ms_c, ms_w, kst = sdpg.bench(data, PARAM, 0, mode, 0, a.at, a.warmup, a.steps,
                                           dump_prefix=dump if idx == 0 else None)
        print(f"[bench] {name}: {a.steps} steps on iterate {a.at}: schur {ms_s:.3f} ms, potrf {ms_c:.3f} ms, "
              f"wall {ms_w:.3f} ms per step; kernels {kst}", flush=True)
        r_out.update({"step_schur_ms": ms_s, "step_chol_ms": ms_c, "step_wall_ms": ms_w,
                      "kernels": {k: {"launches": v[0], "ms_per_step": v[1] / a.steps,
                                      "avg_launch_us": 1e3 * v[1] / max(v[0], 1)} for k, v in kst.items()},
                      "_kst": kst})
        res[name] = r_out

    main_form = "ozaki" if "ozaki" in res else list(res)[0]
    step_ms = res[main_form]["step_wall_ms"]
    out = {
        "metric": "Schur build+Cholesky s/iter @256-bit, gap<=1e-60: Ozaki-MFMA vs plain limb",
        "value": step_ms / 1e3,
        "unit": "s/iter",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": step_ms,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": f"mp{32 * L}",
        "data": "synthetic (seeded max-cut instance from gen_problems.py; no external data)",
        "config": {"workload": f"max-cut n=m={info['m']}, one block, A_i=e_i e_i^T, 256-bit, "
                               f"epsilonStar=epsilonDash=1e-60", "form": main_form, "replay_iteration": a.at,
                   "input_sha256": sha},
    }
    kst = res[main_form]["_kst"]
    oz = kst["ozaki_mfma"] if main_form == "ozaki" else None
    if oz and oz[0]:
        avg_s = oz[1] / 1e3 / oz[0]
        ach = oz[2] / oz[0] / avg_s
        traffic = None
        tfile = os.path.join(ROOT, "profiles", "k_gemm_oz_maxcut1000_traffic.json")
        if os.path.exists(tfile):  # rocprofv3 FETCH_SIZE / WRITE_SIZE passes, copy-calibrated
            with open(tfile) as f:
                traffic = json.load(f)["traffic_bytes_per_launch"]
        out["roofline"] = {"kernel": "k_gemm_oz (Ozaki INT8 slice products, v_mfma_i32_16x16x64_i8)",
                           "bound": "mfma", "achieved": ach / 1e12, "peak": INT8_PEAK / 1e12, "unit": "TFLOP/s",
                           "frac": ach / INT8_PEAK, "traffic": traffic,
                           "ops_per_launch": oz[2] / oz[0], "avg_launch_us": avg_s * 1e6,
                           "note": "INT8 ops: 2 * S(S+1)/2 per mp multiply-add, S = 36 slices; traffic from "
                                   "profiles/k_gemm_oz_maxcut1000_traffic.txt"}
    shares = {k: v[1] / a.steps / step_ms for k, v in kst.items()}
    out["step_breakdown"] = {"share_of_step": shares, "dominant": max(shares, key=shares.get)}
    for f in res.values():
        f.pop("_kst", None)
    out["forms"] = res
    if "plain" in res and "ozaki" in res:
        out["ozaki_speedup_vs_plain"] = res["plain"]["step_wall_ms"] / max(res["ozaki"]["step_wall_ms"], 1e-9)
        if "objP" in res["plain"] and "objP" in res["ozaki"]:
            from decimal import Decimal
            pa, po = Decimal(res["plain"]["objP"]), Decimal(res["ozaki"]["objP"])
            out["objective_rel_diff_ozaki_vs_plain"] = float(abs(pa - po) / max(abs(pa), Decimal(1)))
    if not a.no_cpu and os.path.exists(dump + "X.bin"):
        import oracle_lib

        if not os.path.exists(oracle_lib.ORC_LIB):
            oracle_lib.build_oracle()
        n = min(a.cpu_n, info["m"])
        tb, tc = oracle_lib.time_point(data, 256, dump + "X.bin", dump + "invZ.bin", L, n, 1)
        scale = (info["m"] / n) ** 3
        out["cpu_baseline"] = {"value": tb + tc * scale, "unit": "s/iter", "cores": 1, "kind": "port",
                               "sample": f"oracle (GMP-mpf-semantics C restatement, serial Rpotrf structure) on "
                                         f"the GPU's iterate {a.at}: Make_bMat in full ({tb:.2f} s) and the "
                                         f"Cholesky of B's leading {n}x{n} block ({tc:.2f} s) scaled by "
                                         f"({info['m']}/{n})^3"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
