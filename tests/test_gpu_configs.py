"""Workloads at their real sizes.

Fused beam batch: a seeded subset of the 256-beam batch generator
(scenes.batch_beams(256, seed=3): hex dims U{4..24} x U{4..24} x U{16..96}
at 2.5 mm cells, CoNH/StVK, lognormal E, random clamp face, gravity plus
traction over 10 stages) with its smallest and its largest instance, the
9x6x61 StVK beam whose linear tolerances part ways, and one beam at the
generator's minimum dims 4x4x16.  Each instance solves in the mixed-precision
fused batch to |g| <= 1e-8 (FP64 gradient), equals its solve alone within
1e-6 x its bounding-box diagonal, and the instances the FP64 oracle solves
in seconds equal the oracle's equilibrium within the same bound.

Serving mix: requests drawn from the serving generator at 2k..50k-tet beams
and 32..128-node cloths, plus both range ends, submitted to the request loop
while it runs; every result equals the oracle's (SolveFull,
OptimSolver.cpp:163-264, of Simulable_ThinShell / the tet model) within
1e-6 x bbox.  The oracle solves run in a thread pool beside the GPU (its
ctypes calls release the interpreter lock)."""
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import physim_amd as pg
from oracle import orc
from physim_amd import scenes, workloads

pytestmark = pytest.mark.gpu

OPTS = dict(pg.CONVERGE_OPTS, numMaxIters=200)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_x(scene):
    if hasattr(scene, "hinges"):
        T, kind, l1, l2, rho, extra = scene.oracle_arrays()
        O = orc.Model(scene.X0, T, kind, l1, l2, rho, extra)
    else:
        O = orc.Model(scene.X0, scene.T, scene.material, scene.lame1, scene.lame2,
                      scene.density)
    scene.apply(O)
    r = O.newton(**OPTS)
    return r, O.get_x().reshape(-1, 3)


def _batch_subset():
    scs = scenes.batch_beams(256, seed=3)
    by_size = sorted(range(len(scs)), key=lambda k: scs[k].m)
    pick = [by_size[0], by_size[1], by_size[len(scs) // 2], 32, by_size[-1]]
    assert scs[32].name == "beam(9, 6, 61)_stvk"
    out = [scs[k] for k in pick]
    rng = np.random.default_rng(21)
    out.append(scenes.batch_beam(rng, dmin=(4, 4, 16), dmax=(4, 4, 16)))
    return out


@pytest.mark.parametrize("precond", [0, 2])
def test_beam_batch_real_sizes(ctx, precond):
    """precond 0: the reference's Jacobi; 2: one two-level coarse level per
    instance (the batch workload's setting)."""
    sc = _batch_subset()
    opts = dict(OPTS, precond=precond)
    _say(f"fused beam batch (precond {precond}): "
         + ", ".join(f"{s.name} ({s.m} tets)" for s in sc))
    small = [k for k, s in enumerate(sc) if s.m <= 6000]
    assert len(small) >= 2
    pool = ThreadPoolExecutor(max_workers=4)
    futs = {k: pool.submit(_oracle_x, sc[k]) for k in small}
    B = pg.Batch(ctx, pg.MIXED)
    for s in sc:
        B.add_scene(s)
    B.build()
    t = time.perf_counter()
    res = B.newton(**opts)
    _say(f"  fused solve {time.perf_counter() - t:.2f}s: "
         + ", ".join(f"{r['result']}/{r['iterations']}" for r in res))
    xs = [B.get_x(k).copy() for k in range(len(sc))]
    B.close()
    for k, (s, r) in enumerate(zip(sc, res)):
        assert r["result"] == "SUCCESS" and r["optimality"] <= 1e-8, (s.name, r)
        A = pg.Batch(ctx, pg.MIXED)
        A.add_scene(s)
        A.build()
        ra = A.newton(**opts)[0]
        err = np.abs(A.get_x(0) - xs[k]).max() / s.bbox_diag()
        A.close()
        _say(f"  {s.name}: alone {ra['result']}/{ra['iterations']}, |mix - alone| "
             f"{err:.1e} bbox")
        assert ra["result"] == "SUCCESS" and err <= 1e-6
    for k in small:
        ro, xo = futs[k].result()
        err = np.abs(xs[k] - xo).max() / sc[k].bbox_diag()
        _say(f"  {sc[k].name}: oracle {ro['result']}/{ro['iterations']}, |gpu - oracle| "
             f"{err:.1e} bbox")
        assert ro["result"] == "SUCCESS" and err <= 1e-6
    pool.shutdown()


@pytest.mark.parametrize("precond", [0, 2])
def test_serving_mix_matches_oracle(precond):
    reqs = workloads._serving_mix(8, 50000, 128, seed=5)
    # both ends of the ranges: 2k-tet and ~44k-tet beams, 32- and 128-node cloths
    reqs += [scenes.cantilever(dims=(5, 5, 20), young_sigma=0.3, jitter=0.05, seed=1,
                               name="beam_5x5x20"),
             scenes.cantilever(dims=(13, 13, 52), young_sigma=0.3, jitter=0.05, seed=2,
                               name="beam_13x13x52"),
             scenes.cloth(32, stages=10, vertical=True, name="cloth_32"),
             scenes.cloth(128, stages=10, vertical=True, name="cloth_128")]
    _say("serving mix: " + ", ".join(s.name for s in reqs))
    pool = ThreadPoolExecutor(max_workers=4)
    futs = [pool.submit(_oracle_x, s) for s in reqs]
    S = pg.Server(hbm_budget_bytes=4 << 30, lanes=2, max_batch=4, **dict(OPTS, precond=precond))
    tickets = [None] * len(reqs)

    def producer():
        rng = np.random.default_rng(9)
        for i, s in enumerate(reqs):
            tickets[i] = S.submit(s)
            time.sleep(float(rng.exponential(0.02)))

    th = threading.Thread(target=producer)
    th.start()
    th.join()
    out = [S.wait(t) for t in tickets]
    st = S.stats()
    S.close()
    _say(f"  server {st}")
    for s, (x, r, lat, _), f in zip(reqs, out, futs):
        ro, xo = f.result()
        err = np.abs(x - xo).max() / s.bbox_diag()
        _say(f"  {s.name}: gpu {r['result']}/{r['iterations']} in {lat * 1e3:.0f} ms, oracle "
             f"{ro['result']}/{ro['iterations']}, |gpu - oracle| {err:.1e} bbox")
        assert r["result"] == "SUCCESS" and ro["result"] == "SUCCESS"
        assert err <= 1e-6
    pool.shutdown()
