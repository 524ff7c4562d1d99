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
the Ozaki form (slice count from the autotuner), timed with nothing but the step's own kernels in the
stream; the per-kernel times (events around every product and block-column launch) come from a
second pass over the same number of steps.

roofline: the dominant kernel of the step, the Cholesky block-column kernel k_potrf_col (one launch per
32-column block column: the diagonal block and the panel rows below it eliminated by the division-free
recurrence in a fixed point of signed 29-bit limbs, the diagonal block's elimination repeated in every
workgroup; a serial chain of 32 pivots per launch on the vector ALUs): algorithmic work = the reference's
multiple-precision multiply-adds for that block column, nb^3/6 (Rpotrf2 on the diagonal block) + r nb^2/2
(Rtrsm on the r panel rows), at L^2 = 64 limb multiply-adds each for a 256-bit request (L = 8 limbs of the
request), over its event-timed average launch, against the measured integer-VALU peak (sdpg_valu_peak:
sustained v_mad_u64_u32 + v_addc limb multiply-adds of the whole device).  kernel_rooflines: the short-K
Ozaki product kernel (INT8 peak), the plain limb product kernel (VALU peak) and, on the dense rank-2
workload, the long-K Ozaki kernel (INT8 peak).

also_measured: the GEMM-heavy workload 3 (one block n = 150, m = 1326 dense rank-2 constraints, all F1)
replayed on its first iterate for both forms at 256, 320 and 200 bits (Ozaki slice counts from the
autotuner for each precision), 2 timed steps each.

also_measured_batch: the seeded batch workload (--batch = 512 small SDPs, 1-3 dense blocks of U[8,40],
m in U[10,120]) solved to gap <= 1e-60 in lockstep (every device call of all running problems batched,
finished problems masked out, the Newton phases replayed as a HIP graph): solves/s, latency p50 / p99
from the start of the loops, device ms per phase; the first --batch-independent problems again as
concurrent independent solves (--batch-threads host threads, one context and stream each).

cpu_baseline: the oracle (C restatement of the reference's mpf path: Make_bMat and the reference's
serial Rpotrf, one core as the reference runs them) on the same iterate, dumped by the GPU run: the
whole Make_bMat and the whole 1000 x 1000 Cholesky.  also_measured carries a bounded CPU sample of
workload 3 (F1 precompute with the reference's OpenMP Rgemm on the host's cores, serial inner products).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "sdpa-gmp_amd")
sys.path.insert(0, PKG)
sys.path.insert(0, os.path.join(ROOT, "tests"))

PARAM = os.path.join(PKG, "data", "param_256_gap1e-60.sdpa")
INT8_PEAK = 5.0e15  # MI355X dense INT8 MFMA, ops/s
WORKLOADS = {
    2: "max-cut n=m=1000, one block, A_i=e_i e_i^T, C=-L/4 of G(1000,0.01) with weights +-1",
    3: "one block n=150, m=1326, A_i = u v^T + v u^T (dense), feasible by construction",
    0: "one dense block n=15, m=21",
    1: "blocks (50,25), m=351, about 8% nonzeros per A_i",
}


def structure(path: str) -> dict:
    """m and the block sizes of a problem file"""
    with open(path) as f:
        lines = [ln for ln in f if not ln.startswith('"') and not ln.startswith("*")]
    m = int(lines[0].split()[0])
    nb = int(lines[1].split()[0])
    blocks = [int(x) for x in lines[2].replace("=", " ").split() if x.lstrip("-").isdigit()][:nb]
    return {"m": m, "blocks": blocks}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_cores() -> int:
    """cores this process may use (the box's CPU share), at most 16"""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def measure(workload: int, forms: list, at: int, warmup: int, steps: int, solve: bool, dump: str | None,
            prec: int = 0, data: str | None = None):
    """Replay [Make_bMat + Cholesky] steps of one workload for each GEMM form (after an optional full
    solve); returns (structure, sha256, per-form results)."""
    import gen_problems
    import sdpg

    sha = None
    if data is None:
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
            r = sdpg.solve(data, PARAM, prec, None, mode)
            t_solve = time.time() - t0
            it = max(r.iterations, 1)
            r_out = {"iterations": r.iterations, "phase": r.phase.decode().strip(), "relgap": r.relgap.decode(),
                     "objP": r.objP.decode(), "objD": r.objD.decode(), "solve_wall_s": t_solve,
                     "slices": r.slices,
                     "traj_ms_per_iter": {"schur": r.t_schur_ms / it, "chol": r.t_chol_ms / it,
                                          "solve": r.t_solve_ms / it, "gvec": r.t_gvec_ms / it,
                                          "dxdz": r.t_dxdz_ms / it, "inverse": r.t_inv_ms / it,
                                          "min_eigen": r.t_eig_ms / it}}
            print(f"[bench] w{workload} {name}: solve {r.iterations} iterations, {r_out['phase']}, relgap "
                  f"{r_out['relgap']}, {t_solve:.1f} s; schur+chol {(r.t_schur_ms + r.t_chol_ms) / it:.2f} ms/iter "
                  f"over the trajectory", flush=True)
        ms_s, ms_c, ms_w, kst = sdpg.bench(data, PARAM, prec, mode, 0, at, warmup, steps,
                                           dump_prefix=dump if (idx == 0 and dump) else None)
        print(f"[bench] w{workload} {name} prec {prec or 256}: {steps} steps on iterate {at}: schur {ms_s:.3f} ms, "
              f"potrf {ms_c:.3f} ms, wall {ms_w:.3f} ms per step; kernels {kst}", flush=True)
        r_out.update({"step_schur_ms": ms_s, "step_chol_ms": ms_c, "step_wall_ms": ms_w,
                      "kernels": {k: {"launches": v[0], "ms_per_step": v[1] / steps,
                                      "avg_launch_us": 1e3 * v[1] / max(v[0], 1)} for k, v in kst.items()},
                      "_kst": kst})
        res[name] = r_out
    return info, sha, data, res


def measure_batch(count: int, independent: int, threads: int):
    """The batched configuration: the first `count` problems of the seeded batch (1-3 dense blocks of
    U[8,40], m in U[10,120]) at 256 bits and gap <= 1e-60 solved in lockstep (sdpa_amd_batch_solve: one
    batched call per kind of device call of all running problems, the Newton phases as a HIP graph),
    and the first `independent` of them as concurrent independent solves on `threads` host threads (one
    context and stream each) for comparison on the same problems."""
    import concurrent.futures

    import gen_problems
    import sdpg

    tmp = tempfile.mkdtemp(prefix="sdpg_batch_")
    t0 = time.time()
    paths = gen_problems.generate_batch(tmp, count)
    t_gen = time.time() - t0
    t0 = time.time()
    res, lat, st = sdpg.batch_solve(paths, PARAM, 0, 0, True)
    t_all = time.time() - t0
    wall = st["wall_s"]
    slat = sorted(lat)
    pct = lambda v, q: v[min(len(v) - 1, int(q * len(v)))]  # noqa: E731
    phases = {}
    for r in res:
        ph = r.phase.decode().strip()
        phases[ph] = phases.get(ph, 0) + 1
    out = {"problems": count, "form": "lockstep batched (block kernels over a problem table; every sum of products in "
                                           "block fixed point with exact limb accumulation)",
           "wall_s": wall, "solves_per_s": count / wall, "latency_p50_s": pct(slat, 0.50),
           "latency_p99_s": pct(slat, 0.99), "mean_iterations": sum(r.iterations for r in res) / count,
           "phases": phases, "meetings": st["meetings"], "mean_problems_per_meeting": st["mean_problems_per_meeting"],
           "device_ms_by_phase": st["device_ms"], "setup_and_teardown_s": t_all - wall, "generation_s": t_gen}
    print(f"[bench] batch lockstep: {count} problems: {out['solves_per_s']:.1f} solves/s, p50 "
          f"{out['latency_p50_s']:.3f} s, p99 {out['latency_p99_s']:.3f} s, phases {phases}, device ms "
          f"{ {k: round(v, 1) for k, v in st['device_ms'].items()} }", flush=True)
    if independent > 0:
        sub = paths[:independent]

        def one(path):
            t = time.time()
            r = sdpg.solve(path, PARAM, 0, None, sdpg.GEMM_PLAIN)
            return time.time() - t, r.phase.decode().strip(), r.iterations, r.objP70.decode()

        one(sub[0])  # warm-up: first-use costs of the runtime
        t0 = time.time()
        with concurrent.futures.ThreadPoolExecutor(threads) as ex:
            ind = list(ex.map(one, sub))
        w = time.time() - t0
        il = sorted(x[0] for x in ind)
        same = sum(1 for k, x in enumerate(ind) if x[1] == res[k].phase.decode().strip()
                   and abs(x[2] - res[k].iterations) <= 1)
        out["independent"] = {"problems": independent, "host_threads": threads, "wall_s": w,
                              "solves_per_s": independent / w, "latency_p50_s": pct(il, 0.5),
                              "latency_p99_s": pct(il, 0.99), "same_phase_and_iterations_as_lockstep": same}
        print(f"[bench] batch independent: {independent} problems, {threads} threads: {independent / w:.1f} solves/s",
              flush=True)
    return out


def mfma_roofline(kst_entry, kernel: str, traffic=None, note: str = ""):
    """INT8 MFMA roofline of an Ozaki product kernel from its live event timing (launches, ms, INT8 ops)."""
    n, ms, ops = kst_entry
    if not n:
        return None
    avg_s = ms / 1e3 / n
    ach = ops / n / avg_s
    return {"kernel": kernel, "bound": "mfma", "achieved": ach / 1e12, "peak": INT8_PEAK / 1e12, "unit": "TOP/s",
            "frac": ach / INT8_PEAK, "traffic": traffic, "ops_per_launch": ops / n, "avg_launch_us": avg_s * 1e6,
            "note": note}


def valu_roofline(kst_entry, kernel: str, peak: float, limbs_per_mac: int, traffic=None, note: str = ""):
    """Integer-VALU roofline of a limb-arithmetic kernel: mp multiply-adds x limb multiply-adds each over
    its event-timed average launch, against the measured limb multiply-add peak."""
    n, ms, mp_macs = kst_entry
    if not n:
        return None
    avg_s = ms / 1e3 / n
    ach = mp_macs * limbs_per_mac / n / avg_s
    return {"kernel": kernel, "bound": "valu", "achieved": ach / 1e12, "peak": peak / 1e12,
            "unit": "T limb-MAD/s", "frac": ach / peak, "traffic": traffic,
            "work_per_launch": mp_macs * limbs_per_mac / n, "avg_launch_us": avg_s * 1e6, "note": note}


def load_traffic(name: str):
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)["traffic_bytes_per_launch"]
    return None


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
    ap.add_argument("--extra-precs", default="256,320,200", help="precisions of the second workload's replays")
    ap.add_argument("--cpu-n", type=int, default=0, help="Cholesky sample size of the CPU baseline (0: all of B)")
    ap.add_argument("--batch", type=int, default=512,
                    help="problems of the seeded 512-problem batch solved in lockstep (0: skip)")
    ap.add_argument("--batch-independent", type=int, default=64,
                    help="of those, solved again as concurrent independent solves for comparison (0: skip)")
    ap.add_argument("--batch-threads", type=int, default=16)
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (multi-GPU sharding of the "
                                   "Schur rows and a 2D block-cyclic Cholesky are design notes only)"}))
        sys.exit(2)

    import sdpg

    L = sdpg.limbs_for_bits(256)
    forms = a.forms.split(",")
    peak = sdpg.valu_peak()
    print(f"[bench] integer-VALU peak: {peak / 1e12:.2f} T limb-MAD/s", flush=True)
    tuned = {}
    for prec in sorted({256} | {int(p) for p in a.extra_precs.split(",") if p}):
        S, rows = sdpg.tune_slices(prec)
        tuned[prec] = {"slices": S, "candidates": [{"slices": int(r[0]), "effective_bits": round(r[1], 1),
                                                    "ms_512cube": r[2], "ms_1024sq_k32_lower": r[3]} for r in rows]}
        print(f"[bench] slice autotuner, {prec} bits: S = {S} ({tuned[prec]['candidates']})", flush=True)
    tmp = tempfile.mkdtemp(prefix="sdpg_dump_")
    dump = os.path.join(tmp, "it_")
    info, sha, data2, res = measure(a.workload, forms, a.at, a.warmup, a.steps, not a.no_solve, dump)

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
                   "replay_iteration": a.at, "input_sha256": sha, "ozaki_slices": tuned[256]["slices"],
                   "slice_autotuner": tuned},
    }
    kst = res[main_form]["_kst"]
    shares = {k: v[1] / a.steps / step_ms for k, v in kst.items()}
    dom = max(shares, key=shares.get)
    # the mp multiply-adds of a launch are nb^3/6 + r nb^2/2 (Rpotrf2 on the diagonal block and Rtrsm on the panel
    # rows); 64 limb multiply-adds each for a 256-bit request (8 limbs)
    limbs_req = 8
    rl = valu_roofline(kst["chol_col"], "k_potrf_col (Cholesky 32-column block columns: the diagonal block and the "
                                        "panel rows below it, division-free elimination in signed 29-bit-limb "
                                        "fixed point)", peak, limbs_req * limbs_req,
                       load_traffic("k_potrf_col_maxcut1000_traffic.json"),
                       "work = nb^3/6 + r nb^2/2 mp multiply-adds per block column x 64 limb multiply-adds (256-bit "
                       "request); a serial chain of 32 pivots per launch, one update latency per pivot: "
                       "latency-bound, not throughput-bound")
    if rl:
        out["roofline"] = rl
    kr = {}
    if main_form == "ozaki":
        kr["k_gemm_oz"] = mfma_roofline(kst["ozaki_mfma"], "k_gemm_oz (Ozaki INT8 slice products, short K)",
                                        load_traffic("k_gemm_oz_maxcut1000_traffic.json"),
                                        f"INT8 ops: 2 S(S+1)/2 per mp multiply-add, S = {tuned[256]['slices']}; "
                                        "Cholesky trailing updates in pairs of block columns: the next block column "
                                        "(K = 32), then the trailing matrix (K = 64)")
    if "plain" in res:
        kr["k_gemm_plain"] = valu_roofline(res["plain"]["_kst"]["plain_limb"], "k_gemm_plain (plain limb products)",
                                           peak, limbs_req * limbs_req, None,
                                           "plain form, max-cut Cholesky trailing updates (K = 32)")
    out["step_breakdown"] = {"share_of_step": shares, "dominant": dom}
    for f in res.values():
        f.pop("_kst", None)
    out["forms"] = res
    if "plain" in res and "ozaki" in res:
        out["ozaki_speedup_vs_plain"] = res["plain"]["step_wall_ms"] / max(res["ozaki"]["step_wall_ms"], 1e-9)
        if "objP" in res["plain"] and "objP" in res["ozaki"]:
            from decimal import Decimal
            pa, po = Decimal(res["plain"]["objP"]), Decimal(res["ozaki"]["objP"])
            out["objective_rel_diff_ozaki_vs_plain"] = float(abs(pa - po) / max(abs(pa), Decimal(1)))

    dump3 = os.path.join(tmp, "w3_")
    if not a.no_extra and a.workload != 3:
        # the GEMM-heavy workload: F1 products (A_i Z^-1, X F_i: 2 m n^3) and the Schur inner products
        # (m(m+1)/2 x n(n+1)/2) dominate, the regime where the INT8 form pays
        ex = {"workload": WORKLOADS[3], "replay_iteration": 1, "steps": 2, "by_precision": {}}
        data3 = None
        for prec in [int(p) for p in a.extra_precs.split(",") if p]:
            info3, sha3, data3, res3 = measure(3, forms, 1, 1, 2, False, dump3 if prec == 256 else None,
                                               prec=prec, data=data3)
            if sha3:
                ex["input_sha256"] = sha3
            e = {"slices": tuned[prec]["slices"]}
            for name, r in res3.items():
                e[name] = {"s_per_iter": r["step_wall_ms"] / 1e3, "schur_ms": r["step_schur_ms"],
                           "chol_ms": r["step_chol_ms"], "kernels": r["kernels"]}
            if "ozaki" in res3 and "plain" in res3:
                e["ozaki_speedup_vs_plain"] = res3["plain"]["step_wall_ms"] / max(res3["ozaki"]["step_wall_ms"], 1e-9)
            if "ozaki" in res3:
                S = tuned[prec]["slices"]
                e["roofline_k_gemm_ozk"] = mfma_roofline(
                    res3["ozaki"]["_kst"]["ozaki_mfma_longk"], "k_gemm_ozk (Ozaki INT8 slice products, long K)",
                    load_traffic("k_gemm_ozk_dense_rank2_m1326_traffic.json") if prec == 256 else None,
                    f"INT8 ops: 2 S(S+1)/2 per mp multiply-add, S = {S}; F1 products (K = n) and Schur inner "
                    f"products (K = n(n+1)/2)")
            if "plain" in res3:
                e["roofline_k_gemm_plain"] = valu_roofline(res3["plain"]["_kst"]["plain_limb"],
                                                           "k_gemm_plain (plain limb products)", peak,
                                                           ((prec + 31) // 32) ** 2)
            ex["by_precision"][str(prec)] = e
        out["also_measured"] = ex
    out["kernel_rooflines"] = kr

    if a.batch > 0:
        out["also_measured_batch"] = measure_batch(a.batch, a.batch_independent, a.batch_threads)

    if not a.no_cpu and os.path.exists(dump + "X.bin"):
        import oracle_lib

        if not os.path.exists(oracle_lib.ORC_LIB):
            oracle_lib.build_oracle()
        n = info["m"] if a.cpu_n <= 0 else min(a.cpu_n, info["m"])
        tb, tc = oracle_lib.time_point(data2, 256, dump + "X.bin", dump + "invZ.bin", L, n, 1)
        scale = (info["m"] / n) ** 3
        out["cpu_baseline"] = {"value": tb + tc * scale, "unit": "s/iter", "cores": 1, "kind": "port",
                               "cpu_model": cpu_model(),
                               "sample": f"oracle (GMP-mpf-semantics C restatement of the reference's Make_bMat and "
                                         f"serial Rpotrf; the reference runs both on one core) on the GPU's iterate "
                                         f"{a.at}: Make_bMat in full ({tb:.2f} s) and the Cholesky of B's leading "
                                         f"{n}x{n} block ({tc:.2f} s)" + (f" scaled by ({info['m']}/{n})^3"
                                                                         if n < info["m"] else " (all of B)")}
        if "also_measured" in out and os.path.exists(dump3 + "X.bin"):
            cores = host_cores()
            nsamp = 6
            tcon, tpair = oracle_lib.time_f1_sample(data3, 256, dump3 + "X.bin", dump3 + "invZ.bin", L, nsamp, cores)
            m3 = 1326
            t_schur = m3 * tcon + m3 * (m3 + 1) / 2 * tpair
            t_chol = tc * scale * (m3 / info["m"]) ** 3
            out["also_measured"]["cpu_baseline_256"] = {
                "value": t_schur + t_chol, "unit": "s/iter", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
                "sample": f"oracle on workload 3's iterate 1: F1 precompute of {nsamp} constraints (OpenMP Rgemm NN on "
                          f"{cores} cores, {tcon:.3f} s each) and {nsamp * nsamp} inner products (serial, "
                          f"{tpair * 1e3:.3f} ms each), scaled to 1326 constraints and 879801 pairs ({t_schur:.1f} s); "
                          f"Cholesky from the max-cut measurement scaled by (1326/1000)^3 ({t_chol:.1f} s)"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
