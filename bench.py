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
the Ozaki form, timed with nothing but the step's own kernels in the stream; the per-kernel times
(events around every product and diagonal-block launch) come from a second pass over the same
number of steps.

roofline: the Ozaki MFMA product kernel, INT8 operations = 2 * S(S+1)/2 per mp multiply-add of the
reference loop (S slices), over its event-timed average launch duration in the timed steps, against
the 5.0 POP/s dense INT8 peak.  The per-kernel breakdown of the step is reported beside it, with the
dominant kernel (the Cholesky diagonal-block pivot chain, a latency bound).

also_measured: the GEMM-heavy workload 3 (one block n = 150, m = 1326 dense rank-2 constraints, all F1),
replayed on its first iterate for both forms (2 timed steps), with the roofline of the long-K Ozaki
kernel that carries its F1 products and Schur inner products.

also_measured_batch: the first --batch problems of the seeded batch workload (512 small SDPs, 1-3
dense blocks of U[8,40], m in U[10,120]) solved to gap <= 1e-60 as concurrent independent solves
(--batch-threads host threads, one context and stream each): solves/s and latency p50 / p99.

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
WORKLOADS = {
    2: "max-cut n=m=1000, one block, A_i=e_i e_i^T, C=-L/4 of G(1000,0.01) with weights +-1",
    3: "one block n=150, m=1326, A_i = u v^T + v u^T (dense), feasible by construction",
    0: "one dense block n=15, m=21",
    1: "blocks (50,25), m=351, about 8% nonzeros per A_i",
}


def structure(path: str) -> dict:
    """m, block sizes and the reference's mp multiply-add count of one Make_bMat + Cholesky."""
    with open(path) as f:
        lines = [ln for ln in f if not ln.startswith('"') and not ln.startswith("*")]
    m = int(lines[0].split()[0])
    nb = int(lines[1].split()[0])
    blocks = [int(x) for x in lines[2].replace("=", " ").split() if x.lstrip("-").isdigit()][:nb]
    return {"m": m, "blocks": blocks, "chol_mp_mac": m ** 3 / 6.0}


def measure(workload: int, forms: list, at: int, warmup: int, steps: int, solve: bool, dump: str | None):
    """Replay [Make_bMat + Cholesky] steps of one workload for each GEMM form (after an optional full
    solve); returns (structure, sha256, per-form results)."""
    import gen_problems
    import sdpg

    tmp = tempfile.mkdtemp(prefix="sdpg_bench_")
    data = os.path.join(tmp, f"w{workload}.dat-s")
    sha = gen_problems.generate(workload, data)
    info = structure(data)
    modes = {"ozaki": sdpg.GEMM_OZAKI, "plain": sdpg.GEMM_PLAIN}
    res = {}
    for idx, name in enumerate(forms):
        mode = modes[name]
        r_out = {}
        if solve:
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
            print(f"[bench] w{workload} {name}: solve {r.iterations} iterations, {r_out['phase']}, relgap "
                  f"{r_out['relgap']}, {t_solve:.1f} s; schur+chol {(r.t_schur_ms + r.t_chol_ms) / it:.2f} ms/iter "
                  f"over the trajectory", flush=True)
        ms_s, ms_c, ms_w, kst = sdpg.bench(data, PARAM, 0, mode, 0, at, warmup, steps,
                                           dump_prefix=dump if (idx == 0 and dump) else None)
        print(f"[bench] w{workload} {name}: {steps} steps on iterate {at}: schur {ms_s:.3f} ms, potrf {ms_c:.3f} ms, "
              f"wall {ms_w:.3f} ms per step; kernels {kst}", flush=True)
        r_out.update({"step_schur_ms": ms_s, "step_chol_ms": ms_c, "step_wall_ms": ms_w,
                      "kernels": {k: {"launches": v[0], "ms_per_step": v[1] / steps,
                                      "avg_launch_us": 1e3 * v[1] / max(v[0], 1)} for k, v in kst.items()},
                      "_kst": kst})
        res[name] = r_out
    return info, sha, res


def measure_batch(count: int, threads: int, gemm_name: str):
    """The batched configuration as independent solves: the first `count` problems of the seeded batch
    (1-3 dense blocks of U[8,40], m in U[10,120]) at 256 bits and gap <= 1e-60, `threads` host threads
    each driving its own context and stream, so that the GPU runs several problems' kernels at once."""
    import concurrent.futures

    import gen_problems
    import sdpg

    tmp = tempfile.mkdtemp(prefix="sdpg_batch_")
    t0 = time.time()
    paths = gen_problems.generate_batch(tmp, count)
    t_gen = time.time() - t0
    mode = {"ozaki": sdpg.GEMM_OZAKI, "plain": sdpg.GEMM_PLAIN}[gemm_name]

    def one(path):
        t = time.time()
        r = sdpg.solve(path, PARAM, 0, None, mode)
        return time.time() - t, r.phase.decode().strip(), r.iterations

    one(paths[0])  # warm-up: first-use costs of the runtime
    t0 = time.time()
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        res = list(ex.map(one, paths))
    wall = time.time() - t0
    lat = sorted(x[0] for x in res)
    phases = {}
    for _, ph, _ in res:
        phases[ph] = phases.get(ph, 0) + 1
    pct = lambda q: lat[min(len(lat) - 1, int(q * len(lat)))]
    out = {"problems": count, "host_threads": threads, "form": gemm_name, "wall_s": wall,
           "solves_per_s": count / wall, "latency_p50_s": pct(0.50), "latency_p99_s": pct(0.99),
           "mean_iterations": sum(x[2] for x in res) / count, "phases": phases, "generation_s": t_gen,
           "not_pdopt": [[k, res[k][1], res[k][2]] for k in range(count) if res[k][1] != "pdOPT"]}
    print(f"[bench] batch: {count} problems, {threads} threads: {out['solves_per_s']:.1f} solves/s, "
          f"p50 {out['latency_p50_s']:.3f} s, p99 {out['latency_p99_s']:.3f} s, phases {phases}", flush=True)
    return out


def roofline(kst_entry, kernel: str, traffic=None, note: str = ""):
    """MFMA roofline of one product kernel from its live event timing (launches, ms, INT8 ops)."""
    n, ms, ops = kst_entry
    if not n:
        return None
    avg_s = ms / 1e3 / n
    ach = ops / n / avg_s
    return {"kernel": kernel, "bound": "mfma", "achieved": ach / 1e12, "peak": INT8_PEAK / 1e12, "unit": "TFLOP/s",
            "frac": ach / INT8_PEAK, "traffic": traffic, "ops_per_launch": ops / n, "avg_launch_us": avg_s * 1e6,
            "note": note}


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
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the second measured workload (n=150, m=1326 dense rank-2 constraints)")
    ap.add_argument("--cpu-n", type=int, default=500, help="Cholesky sample size of the CPU baseline")
    ap.add_argument("--batch", type=int, default=128,
                    help="problems of the seeded 512-problem batch solved as concurrent independent solves (0: skip)")
    ap.add_argument("--batch-threads", type=int, default=16)
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (multi-GPU sharding of the "
                                   "Schur rows and a 2D block-cyclic Cholesky are design notes only)"}))
        sys.exit(2)

    import sdpg

    L = sdpg.limbs_for_bits(256)
    forms = a.forms.split(",")
    tmp = tempfile.mkdtemp(prefix="sdpg_dump_")
    dump = os.path.join(tmp, "it_")
    info, sha, res = measure(a.workload, forms, a.at, a.warmup, a.steps, not a.no_solve, dump)

    main_form = "ozaki" if "ozaki" in res else list(res)[0]
    step_ms = res[main_form]["step_wall_ms"]
    wl = WORKLOADS.get(a.workload, f"workload {a.workload}")
    out = {
        "metric": "Schur build+Cholesky s/iter @256-bit, gap<=1e-60: Ozaki-MFMA vs plain limb",
        "value": step_ms / 1e3,
        "unit": "s/iter",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": step_ms,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": f"mp{32 * L}",
        "data": "synthetic (seeded instance from gen_problems.py; no external data)",
        "config": {"workload": f"{wl}, 256-bit, epsilonStar=epsilonDash=1e-60", "form": main_form,
                   "replay_iteration": a.at, "input_sha256": sha},
    }
    kst = res[main_form]["_kst"]
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "k_gemm_oz_maxcut1000_traffic.json")
    if a.workload == 2 and os.path.exists(tfile):  # rocprofv3 FETCH_SIZE / WRITE_SIZE passes, copy-calibrated
        with open(tfile) as f:
            traffic = json.load(f)["traffic_bytes_per_launch"]
    if main_form == "ozaki":
        rl = roofline(kst["ozaki_mfma"], "k_gemm_oz (Ozaki INT8 slice products, short K, v_mfma_i32_16x16x64_i8)",
                      traffic, "INT8 ops: 2 * S(S+1)/2 per mp multiply-add, S = 36 slices; the Cholesky panel and "
                               "trailing-update products (K = 32); traffic from "
                               "profiles/k_gemm_oz_maxcut1000_traffic.txt")
        if rl:
            out["roofline"] = rl
    shares = {k: v[1] / a.steps / step_ms for k, v in kst.items()}
    dom = max(shares, key=shares.get)
    out["step_breakdown"] = {"share_of_step": shares, "dominant": dom}
    if dom == "chol_diag":
        n, ms, work = kst["chol_diag"]
        out["step_breakdown"]["dominant_kernel"] = {
            "kernel": "k_potrf_diag (32 x 32 diagonal blocks of the Cholesky, 8 workgroups each: the pivot "
                      "chain in every one, the inverse's columns split over them)",
            "bound": "VALU issue of one CU per pivot step, a serial chain of m pivots: one fused "
                     "multiply-multiply-subtract in 320-bit limb arithmetic per entry, in rounds of 256 entries",
            "avg_launch_us": 1e3 * ms / max(n, 1), "us_per_pivot": 1e3 * ms / max(n, 1) / 32}
    for f in res.values():
        f.pop("_kst", None)
    out["forms"] = res
    if "plain" in res and "ozaki" in res:
        out["ozaki_speedup_vs_plain"] = res["plain"]["step_wall_ms"] / max(res["ozaki"]["step_wall_ms"], 1e-9)
        if "objP" in res["plain"] and "objP" in res["ozaki"]:
            from decimal import Decimal
            pa, po = Decimal(res["plain"]["objP"]), Decimal(res["ozaki"]["objP"])
            out["objective_rel_diff_ozaki_vs_plain"] = float(abs(pa - po) / max(abs(pa), Decimal(1)))

    if not a.no_extra and a.workload != 3:
        # the GEMM-heavy workload: F1 products (A_i Z^-1, X F_i: 2 m n^3) and the Schur inner products
        # (m(m+1)/2 x n(n+1)/2) dominate, the regime where the INT8 form pays
        info3, sha3, res3 = measure(3, forms, 1, 1, 2, False, None)
        ex = {"workload": WORKLOADS[3] + ", 256-bit", "input_sha256": sha3, "replay_iteration": 1, "steps": 2}
        for name, r in res3.items():
            ex[name] = {"s_per_iter": r["step_wall_ms"] / 1e3, "schur_ms": r["step_schur_ms"],
                        "chol_ms": r["step_chol_ms"], "kernels": r["kernels"]}
        if "ozaki" in res3 and "plain" in res3:
            ex["ozaki_speedup_vs_plain"] = res3["plain"]["step_wall_ms"] / max(res3["ozaki"]["step_wall_ms"], 1e-9)
        if "ozaki" in res3:
            traffic3 = None
            tfile3 = os.path.join(ROOT, "profiles", "k_gemm_ozk_dense_rank2_m1326_traffic.json")
            if os.path.exists(tfile3):
                with open(tfile3) as f:
                    traffic3 = json.load(f)["traffic_bytes_per_launch"]
            rl3 = roofline(res3["ozaki"]["_kst"]["ozaki_mfma_longk"],
                           "k_gemm_ozk (Ozaki INT8 slice products, long K, v_mfma_i32_16x16x64_i8)", traffic3,
                           "INT8 ops: 2 * S(S+1)/2 per mp multiply-add, S = 36; F1 products (K = n) and Schur "
                           "inner products (K = n(n+1)/2); traffic (mean of the three launches) from "
                           "profiles/k_gemm_ozk_dense_rank2_m1326_traffic.txt")
            if rl3:
                ex["roofline"] = rl3
        out["also_measured"] = ex

    if a.batch > 0:
        b = measure_batch(a.batch, a.batch_threads, "ozaki" if "ozaki" in forms else forms[0])
        b["note"] = ("independent solves, one context and stream per host thread; the lockstep batched form "
                     "(persistent fused kernels per Newton phase) is not built")
        out["also_measured_batch"] = b

    if not a.no_cpu and os.path.exists(dump + "X.bin"):
        import gen_problems
        import oracle_lib

        if not os.path.exists(oracle_lib.ORC_LIB):
            oracle_lib.build_oracle()
        data = os.path.join(tmp, f"w{a.workload}.dat-s")
        gen_problems.generate(a.workload, data)
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
