#!/usr/bin/env python3
"""Benchmark of the MI355X Newton inner loop (libpgn).

Workload: the 40x40x160 hex-grid cantilever (0.1 x 0.1 x 0.4 m), 5 tets per
cell = 1,280,000 CoNH tets / 270,641 nodes / 811,923 DoF, per-tet Young's
modulus lognormal(ln 1e5, 0.3), +-5% node jitter, z = 0 face clamped, gravity
ramped over 10 load stages, quasi-static Newton to |g_free| <= 1e-8.
One step = one full Newton solve from the rest state, inputs resident in HBM.
Path timed: PGN_MIXED (FP32 element Hessians + FP32 PCG inside FP64
iterative refinement; FP64 energy, gradient and residual).  Also reported:
the plain all-FP64 GPU path on the same problem (solve time and the
displacement difference, which must stay within 1e-6 x bbox diagonal), the
fused element E+g+H kernel throughput, the roofline of the dominant kernel
(PCG SpMV, timed live with HIP events on the library stream during the timed
solves) and the CPU oracle baseline.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Newton solves/sec + element E/g/H evals/sec at 1e-6 accuracy; "
          "p50/p99 latency")
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(scene, counts, counts_from):
    """Oracle (FP64 CPU restatement of the reference algorithm: OpenMP element
    loops, serial assembly and Jacobi-CG like the reference) on a bounded
    sample of the same workload: one E+g evaluation, one Hessian evaluation +
    assembly and 20 CG iterations of the 1.28M-tet config, scaled by the
    counts of the reference algorithm as the plain FP64 path ran it (Jacobi
    PCG to 1e-9, cap 1000: counts_from says which solve the counts are
    from); plus the 10x10x40 beam's full oracle solve, timed once."""
    from oracle import orc
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    M = orc.Model(scene.X0, scene.T, scene.material, scene.lame1, scene.lame2,
                  scene.density)
    scene.apply(M)
    for _ in range(scene.stages):
        M.step_load()
    t0 = time.perf_counter()
    M.energy()
    g = M.gradient()
    t_eg = time.perf_counter() - t0
    t0 = time.perf_counter()
    rp, ci, v = M.hessian_lower()
    t_h = time.perf_counter() - t0
    its = 20
    t0 = time.perf_counter()
    orc.cg(rp, ci, v, -g, np.zeros_like(g), 1e-30, its)
    t_cg = (time.perf_counter() - t0) / its
    del M, rp, ci, v
    t_solve = (counts["energy_evals"] * t_eg + counts["hessian_evals"] * t_h
               + counts["pcg_iterations"] * t_cg)
    c1, r1 = _cpu_full_solve("config1")
    return {"value": 1.0 / t_solve, "unit": "Newton solves/s", "cores": threads,
            "kind": "port",
            "sample": (f"oracle on the same 1.28M-tet config: E+g {t_eg:.3f}s, "
                       f"Hessian+assembly {t_h:.3f}s, CG iteration {t_cg * 1e3:.1f}ms "
                       f"(20 timed), scaled by the {counts_from}'s "
                       f"{counts['energy_evals']} E+g / {counts['hessian_evals']} H / "
                       f"{counts['pcg_iterations']} CG iterations -> "
                       f"{t_solve:.1f}s per solve"),
            "config1_full_solve_s": c1, "config1_newton_iterations": r1["iterations"],
            "config1_cg_iterations": r1["cg_iterations"]}


def _cpu_full_solve(which):
    """One full oracle Newton solve of the 10x10x40 beam ("config1") or the
    1.28M-tet beam ("config2") with the reference algorithm's linear solver
    (Jacobi CG to 1e-9, cap 1000) and the benchmark's Newton tolerance."""
    import physim_amd as pg
    from oracle import orc
    from physim_amd import scenes
    sc = scenes.config1() if which == "config1" else scenes.config2()
    O = orc.Model(sc.X0, sc.T, sc.material, sc.lame1, sc.lame2, sc.density)
    sc.apply(O)
    t0 = time.perf_counter()
    r = O.newton(**pg.CONVERGE_OPTS)
    return time.perf_counter() - t0, r


def cpu_study(args):
    """The oracle's full solves on the host: the 10x10x40 beam with all host
    threads and with one, median of 5 after a warm-up (each in a child
    process, so that OMP_NUM_THREADS takes effect), and with --cpu-config2
    the 1.28M-tet beam once with all threads.  No GPU is used."""
    import platform
    import subprocess
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {"cpu": platform.processor() or platform.machine(), "host_cpus": os.cpu_count()}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    out["cpu"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    jobs = [("config1", nthreads, 6), ("config1", 1, 6)]
    if args.cpu_config2:
        jobs.append(("config2", nthreads, 1))
    for which, th, reps in jobs:
        env = dict(os.environ, OMP_NUM_THREADS=str(th))
        p = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-solve", which,
                            "--reps", str(reps)], env=env, capture_output=True, text=True,
                           check=True)
        res = json.loads(p.stdout.strip().splitlines()[-1])
        t = res["times"]
        key = f"{which}_{'1thread' if th == 1 else f'{th}threads'}"
        out[key] = {"threads": th, "runs_s": t,
                    "median_s": float(np.median(t[1:] if len(t) > 1 else t)),
                    "newton_iterations": res["iterations"],
                    "cg_iterations": res["cg_iterations"], "result": res["result"]}
        log(f"cpu {key}: {out[key]}")
    print(json.dumps({"cpu_study": out}), flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel`, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel calibration
    recorded there), or None when no summary holds it."""
    path = os.path.join(ROOT, "profiles", "pcg_spmv_beam40x40x160_hbm_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def kernels_only(M, scene, opts):
    """A short, fixed sequence of the hot kernels at the workload's real
    shape, for hardware-counter (rocprofv3 --pmc) runs: the PCG SpMV
    launches of one capped solve, the fused element kernel,
    and the 256 MiB calibration copy whose counters scale the others."""
    M.reset()
    while M.step_load():
        pass
    g = M.gradient()
    _, r = M.linear_solve(-g, None, tol=1e-30, max_iters=50, reg_iters=1,
                          precond=opts["precond"])
    log(f"kernels-only: PCG {r}")
    for _ in range(5):
        M.time_kernel(0, 1)
        M.time_kernel(7, 1)
    log("kernels-only: done")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--precond", type=int, default=2,
                    help="0 Jacobi, 1 3x3 block Jacobi, 2 two-level (default: block "
                         "Jacobi + aggregated rigid-body coarse space)")
    ap.add_argument("--adaptive", type=int, default=1,
                    help="1 (default): Eisenstat-Walker linear tolerance")
    ap.add_argument("--ls-max-iters", type=int, default=2500,
                    help="PCG iteration cap per linear solve of the timed path "
                         "(default 2500; the reference's lsNumMaxIters is 1000, kept "
                         "by the plain FP64 path)")
    ap.add_argument("--sweep", action="store_true",
                    help="also time one solve per (precond, adaptive) variant")
    ap.add_argument("--kernels-only", action="store_true",
                    help="for hardware-counter runs: at full load, one Hessian, one "
                         "PCG solve capped at 50 iterations, 5 E+g+H launches and 5 "
                         "calibration copies; no solve timing, no JSON line")
    ap.add_argument("--workload", default="beam",
                    choices=["beam", "batch", "rods", "serving", "dynamic", "tiny"],
                    help="beam (default): the 1.28M-tet beam; batch: 256 fused beams; "
                         "rods: 4096 helices; serving: Poisson stream of beams and cloths; "
                         "dynamic: implicit time steps of the 1.28M-tet beam")
    ap.add_argument("--count", type=int, default=0,
                    help="instances / rods / requests / time steps of the batch, rods, "
                         "serving and dynamic workloads (0 = the workload's default)")
    ap.add_argument("--rate", type=float, default=100.0,
                    help="serving: Poisson arrival rate (requests/s)")
    ap.add_argument("--rates", default="",
                    help="serving: comma-separated arrival rates, one JSON line each "
                         "(e.g. 50,100,200,500,1000,2000)")
    ap.add_argument("--lanes", type=int, default=4, help="serving: request-loop lanes")
    ap.add_argument("--cohorts", type=int, default=4,
                    help="serving: fused batches resident per lane")
    ap.add_argument("--max-tets", type=int, default=500_000,
                    help="serving: largest beam request (tets)")
    ap.add_argument("--max-cloth", type=int, default=512,
                    help="serving: largest cloth request (nodes per side)")
    ap.add_argument("--group-by-size", type=int, default=1,
                    help="serving: 1 (default) batches of similar footprint, cheapest "
                         "first; 0 arrival order")
    ap.add_argument("--rod-fp32", type=int, default=1,
                    help="rods: 1 (default) bend-twist Hessians in FP32 (PGN_MIXED), 0 FP64")
    ap.add_argument("--rod-density", type=float, default=3.14e-6,
                    help="rods: line density, kg per metre of rod")
    ap.add_argument("--cpu-study", action="store_true",
                    help="time the oracle's full solves on the host (10x10x40 beam with "
                         "all threads and with one, median of 5) and exit; no GPU")
    ap.add_argument("--cpu-config2", action="store_true",
                    help="--cpu-study: also the 1.28M-tet beam's full solve, once")
    ap.add_argument("--cpu-solve", default="", help=argparse.SUPPRESS)
    ap.add_argument("--reps", type=int, default=1, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_solve:
        ts, r = [], None
        for _ in range(args.reps):
            t, r = _cpu_full_solve(args.cpu_solve)
            ts.append(t)
        print(json.dumps({"times": ts, "iterations": r["iterations"],
                          "cg_iterations": r["cg_iterations"], "result": r["result"]}))
        return
    if args.cpu_study:
        cpu_study(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X and "
              f"has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import physim_amd as pg
    from physim_amd import scenes

    if args.workload != "beam":
        from physim_amd import workloads
        out = getattr(workloads, args.workload)(args, log)
        for line in out if isinstance(out, list) else [out]:
            line.update(metric=METRIC, n_gpus=1, steps=line.get("steps", args.steps),
                        warmup=line.get("warmup", args.warmup), higher_is_better=True,
                        vs_baseline=None, data="synthetic")
            print(json.dumps(line), flush=True)
        return

    t0 = time.perf_counter()
    scene = scenes.config2()
    ctx = pg.Context(0)
    opts = dict(pg.CONVERGE_OPTS, precond=args.precond, adaptive_tol=args.adaptive,
                lsNumMaxIters=args.ls_max_iters)
    M = pg.TetModel(ctx, scene.X0, scene.T, scene.material, scene.lame1, scene.lame2,
                    scene.density, pg.MIXED)
    scene.apply(M)
    log(f"setup {time.perf_counter() - t0:.1f}s: {scene.m} tets, {3 * scene.n} DoF, "
        f"{M.device_bytes() / 2**20:.0f} MiB on device")

    if args.kernels_only:
        kernels_only(M, scene, opts)
        M.close()
        ctx.close()
        return

    for w in range(args.warmup):
        M.reset()
        r = M.newton(**opts)
        log(f"warmup {w}: {r}")

    sweep = []
    if args.sweep:
        for pc in (0, 1, 2):
            for ad in (0, 1):
                M.reset()
                r = M.newton(**dict(pg.CONVERGE_OPTS, precond=pc, adaptive_tol=ad))
                r.update(precond=pc, adaptive=ad, err_vs_timed_variant=None)
                sweep.append((r, M.get_x()))
                log(f"sweep precond={pc} adaptive={ad}: {r}")

    M.set_profiling(True)
    results = []
    ctx.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        M.reset()
        r = M.newton(**opts)
        r["wall_s"] = time.perf_counter() - ts
        results.append(r)
        log(f"step {k}: {r}")
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    M.set_profiling(False)
    ok = all(r["result"] == "SUCCESS" and r["optimality"] <= 1e-8 for r in results)
    x_mixed = M.get_x()
    k1_n, k1_ms = M.profile(0)
    k2_n, k2_ms = M.profile(1)
    spmv_ms = k1_ms / max(k1_n, 1)
    spmv_bytes = M.kernel_bytes(8)

    # element E/g/H throughput (fused kernel, model precision = FP32)
    egh_ms = M.time_kernel(0, 20)
    egh_bytes = M.kernel_bytes(0)
    evals = scene.m / (egh_ms * 1e-3)
    copy_ms = M.time_kernel(7, 20)
    copy_gbs = M.kernel_bytes(7) / (copy_ms * 1e-3) / 1e9
    log(f"E/g/H fp32 {egh_ms:.3f} ms ({evals / 1e9:.2f} G evals/s); "
        f"spmv {spmv_ms * 1e3:.1f} us x {k1_n}; copy {copy_gbs:.0f} GB/s")

    fp64 = {}
    if not args.no_fp64:
        F = pg.TetModel(ctx, scene.X0, scene.T, scene.material, scene.lame1, scene.lame2,
                        scene.density, pg.FP64)
        scene.apply(F)
        ts = time.perf_counter()
        rf = F.newton(**dict(pg.CONVERGE_OPTS, precond=0))
        tf = time.perf_counter() - ts
        err = float(np.abs(F.get_x() - x_mixed).max())
        egh64_ms = F.time_kernel(4, 10)
        # the same FP64 path with the mixed path's solver options, to separate
        # the effect of precision from that of the solver choices
        F.reset()
        ts = time.perf_counter()
        rs = F.newton(**opts)
        ts = time.perf_counter() - ts
        fp64_counts = rf
        fp64 = {"fp64_solve_s": tf, "fp64_result": rf["result"],
                "fp64_newton_iterations": rf["iterations"],
                "fp64_pcg_iterations": rf["pcg_iterations"],
                "max_disp_diff_vs_fp64_m": err,
                "disp_tol_m": 1e-6 * scene.bbox_diag(),
                "matches_fp64": bool(err <= 1e-6 * scene.bbox_diag()),
                "speedup_vs_fp64_solve": tf / (elapsed / args.steps),
                "egh_fp64_evals_per_s": scene.m / (egh64_ms * 1e-3),
                "egh_speedup_vs_fp64": egh64_ms / egh_ms,
                "fp64_same_opts_solve_s": ts,
                "speedup_vs_fp64_same_opts": ts / (elapsed / args.steps)}
        ok = ok and fp64["matches_fp64"]
        log(f"fp64 plain path: {rf} in {tf:.2f}s, |dx| {err:.2e}")
        F.close()

    cpu = None
    if not args.no_cpu:
        keys = ("energy_evals", "hessian_evals", "pcg_iterations")
        if fp64:
            counts = {k: int(fp64_counts[k]) for k in keys}
            src = "plain FP64 path (the reference algorithm)"
        else:
            counts = {k: int(np.mean([r[k] for r in results])) for k in keys}
            src = "mixed path (no FP64 run)"
        cpu = cpu_baseline(scene, counts, src)
        log(f"cpu baseline {cpu}")

    lat = np.array([r["wall_s"] for r in results]) * 1e3
    lpr = M.profile(2)[0]  # lanes per block row of the SpMV
    # the PCG's SpMV launch of the variant in use (PGN_PCG and PGN_SELL, as the
    # library reads them)
    variant = os.environ.get("PGN_PCG", "cg1")
    col_bytes = M.profile(5)[0]  # SELL-64 column width in use (0: none)
    if variant not in ("split", "fused") and col_bytes:
        spmv_name = f"cg1_sell_kernel<float, {'true' if col_bytes == 2 else 'false'}>"
        spmv_desc = f"{spmv_name} (SELL-64 SpMV of the PCG, {col_bytes} B columns)"
    else:
        spmv_kernel = {"split": "pcg_k1b_kernel", "fused": "pcg_k1_kernel"}.get(
            variant, "cg1_a_kernel")
        spmv_name = f"{spmv_kernel}<float, {lpr}>"
        spmv_desc = f"{spmv_name} (BSR3 SpMV of the PCG)"
    line = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "Newton solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 Hessian/PCG, f64 energy/gradient/residual",
        "data": "synthetic",
        "config": {"workload": "beam_40x40x160_conh_lognormal: 1,280,000 tets, "
                               "811,923 DoF, 10 load stages, Newton to |g|<=1e-8",
                   "precision": "mixed",
                   "precond": ["jacobi", "block-jacobi", "two-level"][args.precond],
                   "linear_tol": "eisenstat-walker" if args.adaptive else "1e-9",
                   "pcg_max_iters": args.ls_max_iters},
        "converged": ok,
        "newton_iterations": int(np.mean([r["iterations"] for r in results])),
        "pcg_iterations": int(np.mean([r["pcg_iterations"] for r in results])),
        "latency_ms_p50": float(np.percentile(lat, 50)),
        "latency_ms_p99": float(np.percentile(lat, 99)),
        "egh_evals_per_s": evals,
        "egh_gbs": egh_bytes / (egh_ms * 1e-3) / 1e9,
        "copy_gbs": copy_gbs,
        **fp64,
        "pcg_update_avg_us": k2_ms / max(k2_n, 1) * 1e3,
        "roofline": {"bound": "hbm",
                     "kernel": spmv_desc,
                     "achieved": spmv_bytes / (spmv_ms * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": spmv_bytes / (spmv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": measured_traffic(spmv_name),
                     "launches": k1_n, "avg_us": spmv_ms * 1e3,
                     "bytes_per_launch": spmv_bytes},
        "cpu_baseline": cpu,
        "spmv_width_us": {str(l): M.profile(2 + c)[1] for c, l in enumerate((4, 8, 16))},
        "spmv_lanes_per_row": lpr,
    }
    if sweep:
        line["sweep"] = [dict(precond=r["precond"], adaptive=r["adaptive"],
                              result=r["result"], seconds=r["seconds"],
                              newton=r["iterations"], pcg=r["pcg_iterations"],
                              max_dx_vs_timed=float(np.abs(xs - x_mixed).max()))
                         for r, xs in sweep]
    print(json.dumps(line), flush=True)
    M.close()
    ctx.close()


if __name__ == "__main__":
    main()
