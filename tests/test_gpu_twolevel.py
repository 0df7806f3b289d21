"""The two-level preconditioner (precond 2) on the device against the numpy
restatement (oracle/twolevel.py) over the oracle's Hessian of the same x:
the aggregation and kept columns exactly, the coarse operator Ac = P^T H P
and its inverse, and z = M^-1 r; then Newton solves with it against the FP64
oracle's equilibrium (1e-6 x bounding-box diagonal), graph replays against
direct launches, and a change of the fixed DoFs between solves."""
import sys

import numpy as np
import pytest

import physim_amd as pg
from oracle import twolevel as tl
from physim_amd import scenes
from tests.helpers import oracle_model, perturbed_x, pgn_model, small_beam

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dense(O):
    from oracle import orc
    return orc.lower_to_dense(*O.hessian_lower())


@pytest.fixture(scope="module")
def beam():
    return scenes.cantilever(dims=(3, 3, 12), scale=(0.1, 0.1, 0.4), young_sigma=0.3,
                             jitter=0.05, seed=5)


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
def test_coarse_space_matches_oracle(ctx, beam, precision):
    _say(f"twolevel: coarse space precision={precision}")
    s = beam
    # a state close to rest: the Hessian (and so Ac) is positive definite
    x = perturbed_x(s, amp=1e-5, seed=1)
    O = oracle_model(s)
    O.set_x(x)
    A = _dense(O)
    M = pgn_model(ctx, s, precision)
    M.set_x(x)
    agg, cm, Ac, Ai = M.coarse_space(matrices=True)
    want = tl.aggregate(s.X0, 128)
    assert np.array_equal(agg, want)
    assert np.array_equal(cm, tl.kept_columns(s.X0, want, s.fixed))
    P, keep = tl.prolongation(s.X0, want, s.fixed, cm)
    Ac_o = tl.coarse_matrix(A, P, keep)
    tol = 1e-11 if precision == pg.FP64 else 2e-6
    assert np.abs(Ac - Ac_o).max() <= tol * np.abs(Ac_o).max()
    # the inverse of the device's own Ac, compared in the unit-diagonal scaling
    # (rotation and translation columns differ in scale by the squared offsets)
    inv_o = np.linalg.inv(Ac)
    d = np.sqrt(np.diag(Ac))
    dev = (Ai - inv_o) * d[:, None] * d[None, :]
    assert np.abs(dev).max() <= 1e-9 * np.abs(inv_o * d[:, None] * d[None, :]).max()
    assert np.array_equal(Ai, Ai.T)  # exactly symmetric
    rng = np.random.default_rng(3)
    r = rng.standard_normal(3 * s.n)
    r[s.fixed.reshape(-1).astype(bool)] = 0.0
    z = M.precond_apply(r, 2)
    z_o = tl.apply(A, r, P, Ac_o)
    tolz = 1e-9 if precision == pg.FP64 else 1e-4
    assert np.abs(z - z_o).max() <= tolz * np.abs(z_o).max()
    z1 = M.precond_apply(r, 1)
    assert np.abs(z1 - tl.block_jacobi(A, r)).max() <= tolz * np.abs(z1).max()
    M.close()


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
def test_newton_two_level_matches_oracle(ctx, precision):
    _say(f"twolevel: newton vs oracle precision={precision}")
    s = scenes.cantilever(dims=(4, 4, 16), scale=(0.1, 0.1, 0.4), young_sigma=0.3,
                          jitter=0.05, seed=2)
    O = oracle_model(s)
    ro = O.newton(**pg.CONVERGE_OPTS)
    assert ro["result"] == "SUCCESS"
    xo = O.get_x().reshape(-1, 3)
    res = {}
    for precond in (1, 2):
        M = pgn_model(ctx, s, precision)
        r = M.newton(**dict(pg.CONVERGE_OPTS, precond=precond, adaptive_tol=1,
                            lsNumMaxIters=2500))
        _say(f"  precond {precond}: {r}")
        assert r["result"] == "SUCCESS", r
        assert np.abs(M.get_x().reshape(-1, 3) - xo).max() <= 1e-6 * s.bbox_diag()
        res[precond] = r
        M.close()
    assert res[2]["pcg_iterations"] < res[1]["pcg_iterations"]


def test_two_level_graph_replay_matches_direct(ctx, monkeypatch):
    _say("twolevel: graph replay vs direct launches")
    s = scenes.cantilever(dims=(4, 4, 16), scale=(0.1, 0.1, 0.4), young_sigma=0.3,
                          jitter=0.05, seed=2)
    out = []
    for graph in ("0", "1"):
        monkeypatch.setenv("PGN_GRAPH", graph)
        M = pgn_model(ctx, s, pg.MIXED)
        runs = []
        for _ in range(2):
            M.reset()
            r = M.newton(**dict(pg.CONVERGE_OPTS, precond=2, adaptive_tol=1))
            runs.append(((r["result"], r["iterations"], r["pcg_iterations"]), M.get_x().copy()))
        M.close()
        out.append(runs)
    for (rd, xd), (rg, xg) in zip(*out):
        _say(f"  direct {rd} graph {rg}")
        assert rd == rg and rd[0] == "SUCCESS"
        assert np.array_equal(xd, xg)


def test_two_level_after_fixed_dofs_change(ctx):
    _say("twolevel: fixed DoFs changed between solves")
    s = small_beam(dims=(2, 2, 8))
    M = pgn_model(ctx, s, pg.FP64)
    r = M.newton(**dict(pg.CONVERGE_OPTS, precond=2))
    assert r["result"] == "SUCCESS", r
    # clamp the other end as well: the coarse columns of those aggregates change
    fixed = s.fixed.copy()
    fixed[np.abs(s.X0[:, 2] - s.X0[:, 2].max()) < 1e-12] = 1
    s2 = scenes.Scene(s.name + "_both", s.X0, s.T, s.material, s.lame1, s.lame2, s.density,
                      fixed, scenes.G_EARTH, s.stages)
    M.reset()
    s2.apply(M)
    _, cm = M.coarse_space()
    agg = tl.aggregate(s.X0, 128)
    assert np.array_equal(cm, tl.kept_columns(s.X0, agg, fixed))
    r2 = M.newton(**dict(pg.CONVERGE_OPTS, precond=2))
    assert r2["result"] == "SUCCESS", r2
    O = oracle_model(s2)
    assert O.newton(**pg.CONVERGE_OPTS)["result"] == "SUCCESS"
    assert np.abs(M.get_x() - O.get_x()).max() <= 1e-6 * s.bbox_diag()
    M.close()
