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
    st = S.stats()
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
