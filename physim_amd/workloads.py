"""Benchmark workloads beside the 1.28M-tet beam: the 256-instance fused
beam batch, the 4096-rod batch and the serving stream.  Each returns the
fields of one bench.py JSON line (bench.py adds the common ones).

Workloads (all synthetic, seeded):
  batch   -- 256 beams of scenes.batch_beams (hex dims U{4..24} x U{4..24} x
             U{16..96} at 2.5 mm cells, CoNH/StVK 50/50, E lognormal
             (ln 1e5, 0.5), nu U(0.25, 0.45), random clamp face, gravity +
             U(0, 50) Pa traction), one fused pgn_batch solve from rest.
  rods    -- 4096 helices (n U{50..500}, radius U(0.01, 0.05) m, pitch
             U(0.005, 0.05) m, E 1e7, G = E/3, r 1e-3 m), first two vertices
             and first edge angle clamped, gravity over 10 stages, line density
             3.14e-6 kg/m (the configuration leaves it open; heavier rods
             leave more of these very flexible helices unconverged within the
             300-iteration cap of the reference's Newton).
  dynamic -- the 1.28M-tet beam (the quasi-static workload's generator)
             released from rest and integrated by implicit (backward-Euler)
             steps of 1/60 s, gravity ramped over the first 10 steps; one
             step = one Newton solve of the BasicDynamic objective.  The plain
             FP64 path (Jacobi PCG at the reference's tolerances) runs the same
             steps for the speed-up and the displacement check.
  tiny    -- 256 small beams (the batch generator at 1 cm cells, hex dims
             U{2..4} x U{2..4} x U{8..32}) in one fused batch solved with the
             EigenLDLT linear solver (the reference's default: FP64 banded
             LDL^T, one factorization workgroup per instance); the same batch
             with Jacobi-PCG (EigenCG at the reference's tolerances, FP64) is
             timed beside it and the positions compared.
  serving -- Poisson arrivals of a seeded 60/40 mix of beams (the 1.28M-tet
             beam's generator at 2k..max tets) and cloths (n U{32..max}, 1 m
             sheets hanging from the two corners of their top edge, 0.2 kg/m^2),
             through pgn_server; latency = submit -> result on host.
"""
import threading
import time

import numpy as np

import physim_amd as pg
from physim_amd import scenes

SOLVE_OPTS = dict(pg.CONVERGE_OPTS, precond=1, adaptive_tol=1, lsNumMaxIters=2000)
PRECOND_NAMES = ["jacobi", "block-jacobi", "two-level"]


def _opts(args, **kw):
    """SOLVE_OPTS with bench.py's --precond (0 Jacobi, 1 block Jacobi, 2
    two-level; fused batches build one coarse level per instance)."""
    return dict(SOLVE_OPTS, precond=getattr(args, "precond", 1), **kw)


def _stats(vals):
    v = np.asarray(vals, dtype=np.float64)
    return {"min": float(v.min()), "p50": float(np.percentile(v, 50)),
            "max": float(v.max())}


def batch(args, log):
    count = args.count or 256
    t0 = time.perf_counter()
    scs = scenes.batch_beams(count, seed=3)
    ctx = pg.Context(0)
    B = pg.Batch(ctx, pg.MIXED)
    for s in scs:
        B.add_scene(s)
    B.build()
    m = sum(s.m for s in scs)
    n = sum(s.n for s in scs)
    log(f"batch setup {time.perf_counter() - t0:.1f}s: {count} beams, {m} tets, {3 * n} DoF, "
        f"{B.device_bytes() / 2**30:.2f} GiB on device")
    # the reference's own Newton iteration cap (OptimSolver.h:56): buckling
    # StVK instances that stall under rounding-level energy changes run to it;
    # an instance whose linear solve takes over 400 PCG iterations (those
    # buckling ones) continues with block Jacobi (coarse_giveup)
    opts = _opts(args, numMaxIters=100, coarse_giveup=400)
    warm = min(args.warmup, 1)
    for _ in range(warm):
        B.reset()
        B.newton(**opts)
    steps = max(1, min(args.steps, 2))
    times, res = [], None
    for k in range(steps):
        B.reset()
        ctx.synchronize()
        ts = time.perf_counter()
        res = B.newton(**opts)
        times.append(time.perf_counter() - ts)
        ok = sum(r["result"] == "SUCCESS" for r in res)
        log(f"batch step {k}: {times[-1]:.2f}s, {ok}/{count} converged")
    xs = [B.get_x(k).copy() for k in (0, count // 2, count - 1)]
    # the same instances solved alone (each its own model) must agree
    solo, errs = [], []
    for j, k in enumerate((0, count // 2, count - 1)):
        s = scs[k]
        G = pg.TetModel(ctx, s.X0, s.T, s.material, s.lame1, s.lame2, s.density, pg.MIXED)
        s.apply(G)
        ts = time.perf_counter()
        r = G.newton(**opts)
        solo.append(time.perf_counter() - ts)
        errs.append(float(np.abs(G.get_x().reshape(-1, 3) - xs[j]).max() / s.bbox_diag()))
        G.close()
    el = float(np.mean(times))
    B.close()
    ctx.close()
    return {
        "value": count / el, "unit": "Newton solves/s", "ms_per_step": el * 1e3,
        "steps": steps, "warmup": warm, "scaling": "weak",
        "dtype": "f32 Hessian/PCG, f64 energy/gradient/residual",
        "config": {"workload": f"batch_{count}_beams: {m} tets, {3 * n} DoF, one fused "
                               "launch per kernel, per-instance Newton",
                   "precision": "mixed", "precond": PRECOND_NAMES[opts["precond"]],
                   "coarse_giveup": opts["coarse_giveup"],
                   "linear_tol": "eisenstat-walker"},
        "converged": sum(r["result"] == "SUCCESS" for r in res),
        "instances": count,
        "not_converged": {scs[k].name: {"result": r["result"], "iterations": r["iterations"],
                                        "optimality": r["optimality"]}
                          for k, r in enumerate(res) if r["result"] != "SUCCESS"},
        "newton_iterations": _stats([r["iterations"] for r in res]),
        "pcg_iterations_max": max(r["pcg_iterations"] for r in res),
        "solo_check": {"instances": [0, count // 2, count - 1],
                       "max_disp_diff_over_bbox_diag": max(errs),
                       "matches_within_1e-6": bool(max(errs) <= 1e-6),
                       "solo_solve_s": solo},
    }


def tiny(args, log):
    count = args.count or 256
    rng = np.random.default_rng(11)
    scs = [scenes.batch_beam(rng, cell=0.01, dmin=(2, 2, 8), dmax=(4, 4, 32))
           for _ in range(count)]
    m = sum(s.m for s in scs)
    n = sum(s.n for s in scs)
    ctx = pg.Context(0)
    base = dict(pg.CONVERGE_OPTS, lsNumMaxIters=2000, numMaxIters=200)
    steps = max(1, min(args.steps, 3))
    out = {}
    for label, kw in (("ldlt", dict(lsSolverType=pg.LS_EIGENLDLT)),
                      ("cg", dict(lsSolverType=pg.LS_EIGENCG, precond=0))):
        if label == "cg" and args.no_fp64:
            break
        B = pg.Batch(ctx, pg.FP64)
        for s in scs:
            B.add_scene(s)
        B.build()
        for _ in range(min(args.warmup, 1)):
            B.reset()
            B.newton(**base, **kw)
        times = []
        for _ in range(steps if label == "ldlt" else 1):
            B.reset()
            ctx.synchronize()
            t = time.perf_counter()
            res = B.newton(**base, **kw)
            times.append(time.perf_counter() - t)
        out[label] = (float(np.mean(times)), res, [B.get_x(k).copy() for k in range(count)])
        log(f"tiny {label}: {times[-1]:.3f}s, "
            f"{sum(r['result'] == 'SUCCESS' for r in res)}/{count} converged")
        B.close()
    ctx.close()
    el, res, xs = out["ldlt"]
    its = [r["iterations"] for r in res]
    line = {
        "value": count / el, "unit": "Newton solves/s", "ms_per_step": el * 1e3,
        "steps": steps, "warmup": min(args.warmup, 1), "scaling": "weak", "dtype": "f64",
        "config": {"workload": f"tiny_{count}_beams: {m} tets, {3 * n} DoF, EigenLDLT "
                               "(banded LDL^T per instance)", "precision": "fp64"},
        "converged": sum(r["result"] == "SUCCESS" for r in res), "instances": count,
        "newton_iterations": _stats(its),
    }
    if "cg" in out:
        elc, resc, xc = out["cg"]
        both = [k for k in range(count)
                if res[k]["result"] == "SUCCESS" and resc[k]["result"] == "SUCCESS"]
        diff = max((np.abs(xs[k] - xc[k]).max() / scs[k].bbox_diag() for k in both),
                   default=0.0)
        line["eigencg_fp64"] = {"s": elc, "converged":
                                sum(r["result"] == "SUCCESS" for r in resc),
                                "max_disp_diff_over_bbox_diag_both_converged": float(diff)}
    return line


def rods(args, log):
    count = args.count or 4096
    t0 = time.perf_counter()
    sc = scenes.rod_batch(count, seed=4, density=args.rod_density)
    ctx = pg.Context(0)
    prec = pg.MIXED if getattr(args, "rod_fp32", 1) else pg.FP64
    R = pg.Rods(ctx, sc.rods, sc.frame0_normals(), sc.young, sc.shear, sc.radius, sc.radius,
                sc.density, precision=prec)
    R.set_fixed(sc.fixed_mask())
    R.set_gravity([0.0, -9.81, 0.0], sc.stages)
    nv = int(sc.nverts.sum())
    log(f"rods setup {time.perf_counter() - t0:.1f}s: {count} rods, {nv} vertices, "
        f"{R.N} DoF")
    opts = dict(tolMaxError=1e-8, numMaxIters=300, tolMinImpr=0.0, tolMinStep=0.0,
                ls_noise_guard=1e-12)
    warm = min(args.warmup, 1)
    for _ in range(warm):
        R.reset()
        R.newton(**opts)
    steps = max(1, min(args.steps, 3))
    times, res = [], None
    for k in range(steps):
        R.reset()
        ctx.synchronize()
        ts = time.perf_counter()
        res = R.newton(**opts)
        times.append(time.perf_counter() - ts)
        ok = sum(r["result"] == "SUCCESS" for r in res)
        log(f"rods step {k}: {times[-1]:.2f}s, {ok}/{count} converged")
    el = float(np.mean(times))
    R.close()
    ctx.close()
    return {
        "value": count / el, "unit": "Newton solves/s", "ms_per_step": el * 1e3,
        "steps": steps, "warmup": warm, "scaling": "weak",
        "dtype": ("f32 bend-twist Hessians, f64 energy/gradient/banded LDL^T"
                  if prec == pg.MIXED else "f64 elements and banded LDL^T"),
        "config": {"workload": f"rods_{count}_helices: {nv} vertices, line density "
                               f"{args.rod_density:g} kg/m, 10 load stages",
                   "linear_solver": "banded LDL^T per rod"},
        "converged": sum(r["result"] == "SUCCESS" for r in res),
        "instances": count,
        "results": {o: sum(r["result"] == o for r in res)
                    for o in sorted(set(r["result"] for r in res))},
        "newton_iterations": _stats([r["iterations"] for r in res]),
    }


def dynamic(args, log):
    count = args.count or 20
    dt = 1.0 / 60.0
    t0 = time.perf_counter()
    sc = scenes.config2()
    ctx = pg.Context(0)
    M = pg.TetModel(ctx, sc.X0, sc.T, sc.material, sc.lame1, sc.lame2, sc.density, pg.MIXED)
    sc.apply(M)
    log(f"dynamic setup {time.perf_counter() - t0:.1f}s: {sc.m} tets, {3 * sc.n} DoF")
    opts = _opts(args)
    if args.warmup:
        M.dynamic_step(dt, **opts)
        M.reset()
    ctx.synchronize()
    res = []
    ts = time.perf_counter()
    for k in range(count):
        res.append(M.dynamic_step(dt, **opts))
    el = time.perf_counter() - ts
    x_fast = M.get_x()
    log(f"dynamic: {count} steps in {el:.2f}s, newton {[r['iterations'] for r in res]}, "
        f"pcg {[r['pcg_iterations'] for r in res]}")
    out = {
        "value": count / el, "unit": "implicit time steps/s (Newton solves/s)",
        "ms_per_step": el / count * 1e3, "steps": count, "warmup": 1 if args.warmup else 0,
        "scaling": "weak", "dtype": "f32 Hessian/PCG, f64 energy/gradient/residual",
        "config": {"workload": f"dynamic_beam_40x40x160: {sc.m} tets, {3 * sc.n} DoF, "
                               f"backward Euler dt 1/60 s, gravity over 10 steps",
                   "precision": "mixed", "precond": PRECOND_NAMES[opts["precond"]],
                   "linear_tol": "eisenstat-walker"},
        "converged": sum(r["result"] == "SUCCESS" for r in res),
        "newton_iterations": _stats([r["iterations"] for r in res]),
        "pcg_iterations_total": sum(r["pcg_iterations"] for r in res),
    }
    if not args.no_fp64:
        F = pg.TetModel(ctx, sc.X0, sc.T, sc.material, sc.lame1, sc.lame2, sc.density,
                        pg.FP64)
        sc.apply(F)
        ctx.synchronize()
        ts = time.perf_counter()
        rf = [F.dynamic_step(dt, **dict(pg.CONVERGE_OPTS, precond=0)) for _ in range(count)]
        tf = time.perf_counter() - ts
        err = float(np.abs(F.get_x() - x_fast).max())
        out.update(fp64_step_s=tf / count,
                   fp64_converged=sum(r["result"] == "SUCCESS" for r in rf),
                   fp64_pcg_iterations_total=sum(r["pcg_iterations"] for r in rf),
                   speedup_vs_fp64=tf / el, max_disp_diff_vs_fp64_m=err,
                   matches_fp64=bool(err <= 1e-6 * sc.bbox_diag()))
        log(f"dynamic fp64 plain path: {tf:.2f}s, |dx| {err:.2e}")
        F.close()
    M.close()
    ctx.close()
    return out


def _serving_mix(count, max_tets, max_cloth, seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if rng.random() < 0.6:
            m = float(np.exp(rng.uniform(np.log(2000.0), np.log(max_tets))))
            d = max(2, int(round((m / 20.0) ** (1.0 / 3.0))))
            out.append(scenes.cantilever(dims=(d, d, 4 * d), young_sigma=0.3, jitter=0.05,
                                         seed=int(rng.integers(1 << 30)),
                                         name=f"beam_{d}x{d}x{4 * d}"))
        else:
            n = int(rng.integers(32, max_cloth + 1))
            out.append(scenes.cloth(n, stages=10, vertical=True, name=f"cloth_{n}"))
    return out


def _serve_once(reqs, rate, args, log):
    S = pg.Server(hbm_budget_bytes=200 << 30, lanes=args.lanes, max_batch=16,
                  precision=pg.MIXED, group_by_size=args.group_by_size,
                  cohorts_per_lane=args.cohorts, **_opts(args, numMaxIters=300))
    count = len(reqs)
    rng = np.random.default_rng(6)
    gaps = rng.exponential(1.0 / rate, size=count)
    tickets = [None] * count
    t_sub = [0.0] * count

    def producer():
        t = time.perf_counter()
        for i, s in enumerate(reqs):
            t += float(gaps[i])  # the arrival schedule, not the sleep's drift
            d = t - time.perf_counter()
            if d > 0:
                time.sleep(d)
            t_sub[i] = time.perf_counter()
            tickets[i] = S.submit(s)

    t0 = time.perf_counter()
    th = threading.Thread(target=producer)
    th.start()
    th.join()
    lat, queue, done_t, results = [], [], [], []
    for i in range(count):
        x, r, latency, q = S.wait(tickets[i])
        lat.append(latency)
        queue.append(q)
        done_t.append(t_sub[i] + latency)
        results.append(r)
    span = max(done_t) - t0
    offered = max(t_sub) - t0
    st = S.stats()
    S.close()
    ok = sum(r["result"] == "SUCCESS" for r in results)
    thr = count / span
    # saturated: the loop no longer keeps up with the arrivals
    saturated = bool(thr < 0.9 * count / max(offered, 1e-9))
    log(f"serving {rate:g}/s: {ok}/{count} converged in {span:.2f}s ({thr:.1f}/s, "
        f"offered {count / max(offered, 1e-9):.1f}/s{', SATURATED' if saturated else ''}), "
        f"batches {st['batches']}, p50 {np.percentile(lat, 50) * 1e3:.1f} ms "
        f"p99 {np.percentile(lat, 99) * 1e3:.1f} ms")
    return {
        "value": thr, "unit": "Newton solves/s", "ms_per_step": span / count * 1e3,
        "steps": count, "warmup": 0, "scaling": "weak",
        "dtype": "f32 Hessian/PCG, f64 energy/gradient/residual",
        "config": {"workload": f"serving_{count}_requests: Poisson {rate:g}/s, 60% beams "
                               f"2k-{args.max_tets} tets, 40% cloths 32-{args.max_cloth}",
                   "rate": rate, "lanes": args.lanes, "max_batch": 16,
                   "cohorts_per_lane": args.cohorts, "group_by_size": args.group_by_size},
        "offered_per_s": count / max(offered, 1e-9),
        "saturated": saturated,
        "converged": ok,
        "not_converged": {f"{i}:{reqs[i].name}": results[i]["result"] for i in range(count)
                          if results[i]["result"] != "SUCCESS"},
        "latency_ms_p50": float(np.percentile(lat, 50)) * 1e3,
        "latency_ms_p99": float(np.percentile(lat, 99)) * 1e3,
        "queue_ms_p50": float(np.percentile(queue, 50)) * 1e3,
        "queue_ms_p99": float(np.percentile(queue, 99)) * 1e3,
        "batches": st["batches"],
    }


def serving(args, log):
    """Poisson arrivals of the serving mix into the request loop; one result
    per arrival rate of args.rates (comma-separated, else args.rate)."""
    count = args.count or 32
    reqs = _serving_mix(count, args.max_tets, args.max_cloth)
    log(f"serving: {count} requests, {sum(hasattr(r, 'hinges') for r in reqs)} cloths, "
        f"{sum(getattr(r, 'm', 0) for r in reqs)} tets in all")
    rates = [float(v) for v in args.rates.split(",")] if args.rates else [args.rate]
    lines = [_serve_once(reqs, rate, args, log) for rate in rates]
    return lines if args.rates else lines[0]
