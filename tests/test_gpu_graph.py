"""PCG chunks replayed from captured HIP graphs against direct launches: the
same kernels with the same arguments (the chunk's first iteration and the
iteration cap read from device memory), so the same iteration counts and
bit-identical positions.  Every solve replays each chunk graph many times,
with the solver's own direct launches, copies and host synchronizations
between the replays; two solves in a row check that replays after the
first stay exact."""
import sys

import numpy as np
import pytest

import physim_amd as pg
from physim_amd import scenes
from tests.helpers import pgn_model

pytestmark = pytest.mark.gpu

OPTS = dict(pg.CONVERGE_OPTS, precond=1, adaptive_tol=1, lsNumMaxIters=2000, numMaxIters=200)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _solve(ctx, s, precision, graph, monkeypatch, reps=2):
    monkeypatch.setenv("PGN_GRAPH", graph)  # read when the model is created
    M = pgn_model(ctx, s, precision)
    out = []
    for _ in range(reps):
        M.reset()
        r = M.newton(**OPTS)
        out.append(((r["result"], r["iterations"], r["pcg_iterations"]), M.get_x().copy()))
    M.close()
    return out


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
def test_graph_replay_matches_direct_launches(ctx, precision, monkeypatch):
    _say(f"graph replay single model precision={precision}")
    s = scenes.cantilever(dims=(4, 4, 16), scale=(0.1, 0.1, 0.4), young_sigma=0.3,
                          jitter=0.05, seed=2)
    direct = _solve(ctx, s, precision, "0", monkeypatch)
    graph = _solve(ctx, s, precision, "1", monkeypatch)
    for (rd, xd), (rg, xg) in zip(direct, graph):
        _say(f"  direct {rd} graph {rg}")
        assert rd == rg and rd[0] == "SUCCESS"
        assert np.array_equal(xd, xg)


def test_batch_graph_replay_matches_direct_launches(ctx, monkeypatch):
    _say("graph replay batch")
    rng = np.random.default_rng(9)
    sc = [scenes.batch_beam(rng, cell=0.01, dmin=(2, 2, 4), dmax=(4, 4, 10)) for _ in range(6)]
    out = []
    for graph in ("0", "2"):
        monkeypatch.setenv("PGN_GRAPH", graph)
        B = pg.Batch(ctx, pg.MIXED)
        for s in sc:
            B.add_scene(s)
        B.build()
        runs = []
        for _ in range(2):
            B.reset()
            res = B.newton(**OPTS)
            runs.append(([(r["result"], r["pcg_iterations"]) for r in res],
                         [B.get_x(k).copy() for k in range(len(sc))]))
        B.close()
        out.append(runs)
    for (cd, xd), (cg, xg) in zip(out[0], out[1]):
        assert cd == cg
        assert all(np.array_equal(a, b) for a, b in zip(xd, xg))
