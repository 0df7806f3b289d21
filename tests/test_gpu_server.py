"""The request loop (pgn_server) on the GPU: requests submitted while others
run are admitted by footprint, solved in fused batches on several lanes, and
each result matches the FP64 oracle's solo solve within 1e-6 x its
bounding-box diagonal."""
import sys
import threading
import time

import numpy as np
import pytest

import physim_amd as pg
from oracle import orc
from physim_amd import scenes

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
    assert r["result"] == "SUCCESS"
    return O.get_x().reshape(-1, 3)


def _stats_drained(S):
    """stats once the lanes have released every finished batch: a request is
    handed back before its batch's memory is freed"""
    t0 = time.time()
    while True:
        st = S.stats()
        if st["bytes_in_use"] == 0 or time.time() - t0 > 10:
            return st
        time.sleep(0.001)


def _requests(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if i % 3 == 2:
            out.append(scenes.cloth(int(rng.integers(4, 8)), size=0.2, stages=10))
        else:
            out.append(scenes.batch_beam(rng, cell=0.01, dmin=(2, 2, 4), dmax=(4, 4, 10)))
    return out


@pytest.mark.parametrize("group_by_size", [1, 0])
def test_server_results_match_oracle(group_by_size):
    _say(f"server: staggered submissions, 2 lanes, footprint admission, "
         f"group_by_size={group_by_size}")
    reqs = _requests(11, 9)
    # a budget that holds two to five of these requests (0.06-0.5 MB each)
    S = pg.Server(hbm_budget_bytes=1_000_000, lanes=2, max_batch=4,
                  group_by_size=group_by_size, **OPTS)
    tickets = [None] * len(reqs)

    def producer():
        rng = np.random.default_rng(5)
        for i, s in enumerate(reqs):
            tickets[i] = S.submit(s)
            time.sleep(float(rng.exponential(0.005)))

    th = threading.Thread(target=producer)
    th.start()
    th.join()
    lat = []
    for i, s in enumerate(reqs):
        x, r, latency, queue = S.wait(tickets[i])
        assert r["result"] == "SUCCESS", (i, r)
        err = np.abs(x - _oracle_x(s)).max()
        assert err <= 1e-6 * s.bbox_diag(), (i, err)
        assert 0.0 <= queue <= latency
        lat.append(latency)
    st = _stats_drained(S)
    _say(f"  stats {st}, latency p50 {np.percentile(lat, 50):.3f}s "
         f"p99 {np.percentile(lat, 99):.3f}s")
    assert st["solved"] == len(reqs) and st["bytes_in_use"] == 0
    assert st["batches"] >= 3  # the budget forced several admissions
    S.close()


def test_server_rejects_request_over_budget():
    _say("server: over-budget request")
    S = pg.Server(hbm_budget_bytes=1000, lanes=1, max_batch=2, **OPTS)
    s = scenes.cantilever(dims=(2, 2, 4))
    with pytest.raises(pg.PgnError, match="budget"):
        S.submit(s)
    S.close()


def _gpu_solo(ctx, scene):
    B = pg.Batch(ctx, pg.MIXED)
    B.add_scene(scene)
    B.build()
    r = B.newton(**OPTS)
    assert r[0]["result"] == "SUCCESS"
    x = B.get_x(0).copy()
    B.close()
    return x


def test_server_admits_mid_run_and_returns_early(ctx):
    """Continuous batching: with one lane busy on a long solve, a small
    request submitted meanwhile is admitted between Newton iterations, ends
    (and is handed back) while the long one still runs, and equals its solo
    solve within 1e-6 x its bounding-box diagonal."""
    _say("server: a small request admitted while a long one runs (1 lane)")
    big = scenes.config1()  # 20,000 tets over 10 load stages
    small = scenes.cantilever(dims=(2, 2, 4), scale=(0.02, 0.02, 0.04), stages=2)
    S = pg.Server(lanes=1, max_batch=4, cohorts_per_lane=4, **OPTS)
    tb = S.submit(big)
    t0 = time.time()
    while S.stats()["batches"] < 1:  # the long request is on the GPU
        assert time.time() - t0 < 60
        time.sleep(0.001)
    ts = S.submit(small)
    xs, rs, lat_s, _ = S.wait(ts)
    big_done_then = S.poll(tb)
    xb, rb, lat_b, _ = S.wait(tb)
    st = _stats_drained(S)
    S.close()
    _say(f"  small {rs['iterations']} its {lat_s * 1e3:.1f} ms, big {rb['iterations']} its "
         f"{lat_b * 1e3:.1f} ms, big done when small returned: {big_done_then}, {st}")
    assert rs["result"] == "SUCCESS" and rb["result"] == "SUCCESS"
    assert not big_done_then
    assert st["batches"] == 2 and st["solved"] == 2 and st["bytes_in_use"] == 0
    d = small.bbox_diag()
    assert np.abs(xs - _gpu_solo(ctx, small)).max() <= 1e-6 * d
    assert np.abs(xs - _oracle_x(small)).max() <= 1e-6 * d
    assert np.abs(xb - _gpu_solo(ctx, big)).max() <= 1e-6 * big.bbox_diag()


@pytest.mark.parametrize("cohorts", [1, 3])
def test_server_cohort_limit(cohorts):
    """cohorts_per_lane bounds the fused batches resident on a lane; results
    do not depend on it."""
    _say(f"server: cohorts_per_lane={cohorts}")
    reqs = _requests(17, 6)
    S = pg.Server(lanes=1, max_batch=1, cohorts_per_lane=cohorts, **OPTS)
    tickets = [S.submit(s) for s in reqs]
    for s, t in zip(reqs, tickets):
        x, r, _, _ = S.wait(t)
        assert r["result"] == "SUCCESS"
        assert np.abs(x - _oracle_x(s)).max() <= 1e-6 * s.bbox_diag()
    st = _stats_drained(S)
    S.close()
    assert st["batches"] == len(reqs) and st["solved"] == len(reqs)


def test_server_destroy_during_a_solve_ends_promptly():
    """Destroying the server stops its lanes after their current Newton
    iteration instead of running the resident solves to the end."""
    _say("server: destroy while a long solve runs")
    big = scenes.cantilever(dims=(16, 16, 64), scale=(0.1, 0.1, 0.4), stages=20)  # 81,920 tets
    S = pg.Server(lanes=1, **OPTS)
    t0 = time.perf_counter()
    _, r, _, _ = S.wait(S.submit(big))
    t_full = time.perf_counter() - t0
    S.close()
    assert r["result"] == "SUCCESS"
    S = pg.Server(lanes=1, **OPTS)
    S.submit(big)
    time.sleep(0.3 * t_full)
    t1 = time.perf_counter()
    S.close()
    dt = time.perf_counter() - t1
    _say(f"  full solve {t_full * 1e3:.0f} ms ({r['iterations']} its), destroy after "
         f"{0.3 * t_full * 1e3:.0f} ms took {dt * 1e3:.1f} ms")
    assert dt < 0.3 * t_full
