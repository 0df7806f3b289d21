"""libpgn's EigenLDLT linear solver (FP64 banded LDL^T on the device,
LinearSolver_EigenLDLT.cpp:49-153) against the oracle's restatement and
dense linear algebra.

Tolerances: a direct solve of the assembled matrix against numpy's dense
solve of the same exported matrix, relative 1e-10; against the oracle's
solve of its own FP64 Hessian, relative 1e-9 (FP64 models); Newton
positions within 1e-6 x bounding-box diagonal of the oracle's Newton with
EigenLDLT, and the same number of Newton iterations within one.
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import physim_amd as pg
from oracle import orc
from physim_amd import scenes
from tests.helpers import oracle_model, pgn_model, small_beam

pytestmark = pytest.mark.gpu

SOLVE = dict(tolMaxError=1e-9, lsTolMaxError=1e-9, tolMinStep=0.0, tolMinImpr=0.0,
             ls_noise_guard=1e-12)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dense(G):
    rp, ci, v = G.hessian(pg.LAYOUT_BSR3)
    return sp.bsr_matrix((v.reshape(-1, 3, 3), ci, rp), shape=(G.N, G.N)).toarray()


def _loaded(G):
    while G.step_load():
        pass


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
def test_linear_solve_matches_dense(ctx, precision):
    _say(f"ldlt linear solve precision={precision}")
    s = small_beam(dims=(2, 2, 6), seed=4)
    G = pgn_model(ctx, s, precision)
    _loaded(G)
    b = -G.gradient()
    A = _dense(G)
    x, r = G.linear_solve(b, None, tol=1e-9, reg_iters=10, solver=pg.LS_EIGENLDLT)
    assert r["status"] == 0 and r["reg_used"] == 0 and r["iterations"] == 0, r
    xd = np.linalg.solve(A, b)
    assert np.abs(x - xd).max() <= 1e-10 * np.abs(xd).max()
    assert r["rel_residual"] <= 1e-9
    if precision == pg.FP64:
        O = oracle_model(s)
        _loaded(O)
        rp, ci, v = O.hessian_lower()
        xo = np.zeros(G.N)
        st, reg = orc.sparse_ldlt_solve(rp, ci, v, -O.gradient(), xo, 1e-9, 10)
        assert st == 0 and reg == 0
        assert np.abs(x - xo).max() <= 1e-9 * np.abs(xo).max()
    G.close()


def test_regularization_tries_match_oracle(ctx):
    """A StVK beam squashed to 40% of its length: the Hessian is indefinite
    and the tries, their count and the regularized solution follow the
    oracle's restatement on the same (exported) matrix."""
    _say("ldlt regularization")
    s = small_beam(dims=(2, 2, 5), material=1, seed=2)
    s.fixed = np.zeros_like(s.fixed)
    G = pgn_model(ctx, s, pg.FP64)
    x = s.X0.copy()
    x[:, 2] *= 0.4
    G.set_x(x)
    b = -G.gradient()
    rp, ci, v = G.hessian(pg.LAYOUT_LOWER)
    xo = np.zeros(G.N)
    st, reg = orc.sparse_ldlt_solve(rp, ci, v, b, xo, 1e-9, 10)
    xg, r = G.linear_solve(b, None, tol=1e-9, reg_iters=10, solver=pg.LS_EIGENLDLT)
    _say(f"  oracle status {st} reg {reg}; libpgn {r}")
    assert reg > 0  # the case exercises the tries
    assert r["status"] == st and r["reg_used"] == reg
    assert np.abs(xg - xo).max() <= 1e-8 * np.abs(xo).max()
    G.close()


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
@pytest.mark.parametrize("material", [0, 1])
def test_newton_matches_oracle(ctx, precision, material):
    _say(f"ldlt newton material={material} precision={precision}")
    s = small_beam(dims=(2, 3, 7), material=material, seed=5)
    O = oracle_model(s)
    ro = O.newton(**SOLVE, ls_solver=2)
    G = pgn_model(ctx, s, precision)
    rg = G.newton(**SOLVE, lsSolverType=pg.LS_EIGENLDLT)
    _say(f"  oracle {ro}\n  libpgn {rg}")
    assert ro["result"] == "SUCCESS" and rg["result"] == "SUCCESS"
    assert rg["pcg_iterations"] == 0
    assert abs(rg["iterations"] - ro["iterations"]) <= 1
    assert np.abs(G.get_x() - O.get_x()).max() <= 1e-6 * s.bbox_diag()
    G.close()


def test_cloth_newton_matches_oracle(ctx):
    _say("ldlt cloth newton")
    s = scenes.cloth(7, stages=3, vertical=True, seed=3)
    T, kind, l1, l2, rho, extra = s.oracle_arrays()
    O = orc.Model(s.X0, T, kind, l1, l2, rho, extra)
    s.apply(O)
    ro = O.newton(**SOLVE, ls_solver=2)
    G = pg.ShellModel(ctx, s.X0, s.F, s.lame1, s.lame2, s.thickness, s.density, s.hinges,
                      s.bending_k, pg.FP64)
    s.apply(G)
    rg = G.newton(**SOLVE, lsSolverType=pg.LS_EIGENLDLT)
    _say(f"  oracle {ro}\n  libpgn {rg}")
    assert ro["result"] == "SUCCESS" and rg["result"] == "SUCCESS"
    assert np.abs(G.get_x() - O.get_x()).max() <= 1e-6 * s.bbox_diag()
    G.close()


def test_dynamic_steps_match_oracle(ctx):
    _say("ldlt dynamic steps")
    s = small_beam(dims=(2, 2, 6), seed=21)
    s.stages = 3
    O = oracle_model(s)
    D = orc.Dynamic(O)
    G = pgn_model(ctx, s, pg.FP64)
    for k in range(3):
        ro = D.step(0.01, **SOLVE, ls_solver=2)
        rg = G.dynamic_step(0.01, **SOLVE, lsSolverType=pg.LS_EIGENLDLT)
        assert ro["result"] == "SUCCESS" and rg["result"] == "SUCCESS", (k, ro, rg)
        assert np.abs(G.get_x() - O.get_x()).max() <= 1e-6 * s.bbox_diag(), k
    G.close()


def test_band_beyond_limit_is_rejected(ctx):
    _say("ldlt band limit")
    s = scenes.config1()  # 11 x 11 nodes per cross-section: half-bandwidth ~400 DoFs
    G = pgn_model(ctx, s, pg.FP64)
    x0 = G.get_x()
    with pytest.raises(pg.PgnError):
        G.newton(**SOLVE, lsSolverType=pg.LS_EIGENLDLT)
    assert np.array_equal(G.get_x(), x0)
    with pytest.raises(pg.PgnError):  # unknown solver type
        G.newton(**SOLVE, lsSolverType=0)
    G.close()


def test_zero_hessian_fails_and_keeps_x(ctx):
    _say("ldlt zero Hessian")
    X0 = np.random.default_rng(1).random((4, 3))
    G = pg.TetModel(ctx, X0, np.zeros((0, 4), np.int32), np.zeros(0, np.uint8), np.zeros(0),
                    np.zeros(0), np.zeros(0), pg.FP64)
    x0 = np.arange(12.0)
    x, r = G.linear_solve(np.ones(12), x0, tol=1e-9, reg_iters=10, solver=pg.LS_EIGENLDLT)
    assert r["status"] == 1, r
    assert np.array_equal(x, x0)
    G.close()


def test_batch_matches_single_and_oracle(ctx):
    """A ragged batch (one tet, a beam, a cloth) solved with EigenLDLT, one
    factorization workgroup per instance: each instance as alone and as the
    oracle's EigenLDLT Newton."""
    _say("ldlt batch")
    rng = np.random.default_rng(7)
    X0 = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.1, 0.0], [0.02, 0.03, 0.1]])
    l1, l2 = orc.lame(1e5, 0.3)
    fixed = np.zeros((4, 3), dtype=np.uint8)
    fixed[:3] = 1
    tet = scenes.Scene("one_tet", X0, np.array([[0, 1, 2, 3]], dtype=np.int32),
                       np.array([0], dtype=np.uint8), np.array([l1]), np.array([l2]),
                       np.array([1000.0]), fixed, scenes.G_EARTH, 5)
    mix = [tet, scenes.batch_beam(rng, cell=0.01, dmin=(2, 2, 4), dmax=(3, 3, 6)),
           scenes.cloth(5, size=0.2, stages=10)]
    opts = dict(SOLVE, numMaxIters=200, lsSolverType=pg.LS_EIGENLDLT)
    B = pg.Batch(ctx, pg.FP64)
    for s in mix:
        B.add_scene(s)
    B.build()
    res = B.newton(**opts)
    _say(f"  {res}")
    for k, s in enumerate(mix):
        assert res[k]["result"] == "SUCCESS" and res[k]["pcg_iterations"] == 0, (k, res[k])
        A = pg.Batch(ctx, pg.FP64)
        A.add_scene(s)
        A.build()
        ra = A.newton(**opts)
        assert ra[0]["result"] == "SUCCESS", (k, ra)
        assert np.abs(A.get_x(0) - B.get_x(k)).max() <= 1e-9 * s.bbox_diag(), k
        A.close()
    for k in (0, 1):
        O = oracle_model(mix[k])
        ro = O.newton(**SOLVE, numMaxIters=200, ls_solver=2)
        assert ro["result"] == "SUCCESS", (k, ro)
        assert np.abs(B.get_x(k) - O.get_x().reshape(-1, 3)).max() <= 1e-6 * mix[k].bbox_diag()
    B.close()
