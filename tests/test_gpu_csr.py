"""pgn_csr_linear_solve: ILinearSolver::Solve(A, b, x) (LinearSolver.h:38-52,
LinearSolver_EigenCG.cpp:51-125) on a caller's CSR matrix, against the
oracle's restatement (orc_linear_solve: Eigen's Jacobi CG and the cumulative
regularization loop on a lower CSR matrix).

Tolerances: at tol 1e-8 the iteration counts agree within 2 (the first
iterates agree to 1e-15; the CG recurrences round differently: summation
order, fused multiply-adds); at tol 1e-10, near the rounding floor of these
systems (condition ~3e4), the counts may part by a few iterations (109
against 115 on the 2x2x6 beam), so the solution is checked by its true
residual |b - A x| <= tol |b| in FP64 on the host and against the oracle's
to 1e-6 relative; regularization tries and the shifted diagonal are exact
(the same sequence of FP64 additions)."""
import ctypes
import sys

import numpy as np
import pytest
import scipy.sparse

import physim_amd as pg
from oracle import orc
from physim_amd import scenes
from tests.helpers import oracle_model, perturbed_x, pgn_model

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_solve(rp, ci, v, b, x0, tol, maxit, reg):
    x = np.array(x0, dtype=np.float64)
    v = np.array(v, dtype=np.float64)
    it = ctypes.c_int(0)
    used = ctypes.c_int(0)
    st = orc.lib().orc_linear_solve(rp.size - 1, orc._p(rp, orc._lp), orc._p(ci, orc._ip),
                                    orc._p(v), orc._p(np.asarray(b, np.float64)), orc._p(x),
                                    tol, maxit, reg, ctypes.byref(it), ctypes.byref(used))
    return x, {"status": st, "iterations": it.value, "reg_used": used.value}, v


def _lower_to_full(rp, ci, v, n):
    L = scipy.sparse.csr_matrix((v, ci, rp), shape=(n, n))
    A = (scipy.sparse.tril(L) + scipy.sparse.tril(L, -1).T).tocsr()
    A.sort_indices()
    return A


def _beam_hessian(dims=(10, 10, 40), amp=1e-4):
    s = scenes.cantilever(dims=dims, scale=(0.1, 0.1, 0.4))
    O = oracle_model(s)
    for _ in range(s.stages):
        O.step_load()
    O.set_x(perturbed_x(s, amp))
    rp, ci, v = O.hessian_lower()
    return s, O, rp, ci, v


@pytest.mark.parametrize("dims", [(2, 2, 6), (10, 10, 40)])
def test_csr_solve_lower_matches_oracle(ctx, dims):
    _say(f"csr solve lower vs oracle dims={dims}")
    s, O, rp, ci, v = _beam_hessian(dims)
    n = rp.size - 1
    b = np.random.default_rng(5).standard_normal(n)
    x8, r8 = pg.csr_linear_solve(ctx, rp, ci, v.copy(), b, tol=1e-8, max_iters=5000)
    xo8, ro8, _ = _oracle_solve(rp, ci, v, b, np.zeros(n), 1e-8, 5000, 10)
    _say(f"  tol 1e-8: gpu {r8}  oracle {ro8}")
    assert r8["status"] == ro8["status"] == 0 and r8["reg_used"] == ro8["reg_used"]
    assert abs(r8["iterations"] - ro8["iterations"]) <= 2
    assert np.linalg.norm(x8 - xo8) <= 1e-6 * np.linalg.norm(xo8)
    opts = dict(tol=1e-10, max_iters=5000, reg_iters=10, precond=0)
    vg = v.copy()
    x, r = pg.csr_linear_solve(ctx, rp, ci, vg, b, **opts)
    xo, ro, vo = _oracle_solve(rp, ci, v, b, np.zeros(n), 1e-10, 5000, 10)
    _say(f"  tol 1e-10: gpu {r}  oracle {ro}")
    assert r["status"] == ro["status"] == 0
    assert r["reg_used"] == ro["reg_used"]
    A = _lower_to_full(rp, ci, v, n)
    assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b) * 1.01
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    assert np.array_equal(vg, vo)  # untouched without regularization
    # the same matrix given in full runs the same products: identical bits
    F = _lower_to_full(rp, ci, v, n)
    vf = F.data.copy()
    xf, rf = pg.csr_linear_solve(ctx, F.indptr, F.indices, vf, b, layout=pg.LAYOUT_FULL, **opts)
    assert rf == r and np.array_equal(xf, x)


def test_csr_solve_regularizes_like_the_oracle(ctx):
    """diag(4, 3, 2, 1, -0.5, -1), b = e4 + e5: b.x < 0 until the cumulative
    shift regT (10^i - 1)/9 passes 1; every try's shift lands on the caller's
    diagonal."""
    _say("csr solve regularization")
    n = 6
    rp = np.arange(n + 1, dtype=np.int64)
    ci = np.arange(n, dtype=np.int32)
    v = np.array([4.0, 3.0, 2.0, 1.0, -0.5, -1.0])
    b = np.array([0.0, 0.0, 0.0, 0.0, 1.0, 1.0])
    vg = v.copy()
    x, r = pg.csr_linear_solve(ctx, rp, ci, vg, b, tol=1e-12, max_iters=100, reg_iters=10)
    xo, ro, vo = _oracle_solve(rp, ci, v, b, np.zeros(n), 1e-12, 100, 10)
    _say(f"  gpu {r}  oracle {ro}")
    assert r["status"] == ro["status"] == 0 and r["reg_used"] == ro["reg_used"] > 0
    assert np.array_equal(vg, vo)
    np.testing.assert_allclose(x, xo, rtol=1e-12)
    assert b @ x >= 0
    # a system no try rescues fails after reg_iters tries, keeping every shift
    v2 = np.array([1.0, 1.0, 1.0, 1.0, -1e9, -1e9])
    v2g = v2.copy()
    _, r2 = pg.csr_linear_solve(ctx, rp, ci, v2g, b, tol=1e-12, max_iters=100, reg_iters=3)
    _, ro2, vo2 = _oracle_solve(rp, ci, v2, b, np.zeros(n), 1e-12, 100, 3)
    assert r2["status"] == ro2["status"] == 1 and r2["reg_used"] == ro2["reg_used"] == 2
    assert np.array_equal(v2g, vo2)


def test_csr_solve_edge_cases(ctx):
    _say("csr solve edge cases")
    s, O, rp, ci, v = _beam_hessian((2, 2, 4))
    n = rp.size - 1
    # b = 0 -> x = 0, 0 iterations (Eigen)
    x, r = pg.csr_linear_solve(ctx, rp, ci, v.copy(), np.zeros(n), x0=np.ones(n))
    assert r["status"] == 0 and r["iterations"] == 0 and np.all(x == 0.0)
    # a warm start at the solution returns at once
    A = _lower_to_full(rp, ci, v, n)
    xs = np.linalg.solve(A.toarray(), np.ones(n))
    x, r = pg.csr_linear_solve(ctx, rp, ci, v.copy(), np.ones(n), x0=xs, tol=1e-9)
    assert r["iterations"] == 0 and np.array_equal(x, xs)
    # max_iters = 0: no iteration, x kept
    x, r = pg.csr_linear_solve(ctx, rp, ci, v.copy(), np.ones(n), x0=np.zeros(n), max_iters=0)
    assert r["iterations"] == 0 and np.all(x == 0.0)
    # an empty system
    x, r = pg.csr_linear_solve(ctx, np.zeros(1, np.int64), np.zeros(0, np.int32), np.zeros(0),
                               np.zeros(0))
    assert r["status"] == 0 and x.size == 0
    # rejected inputs
    bad = ci.copy()
    bad[3] = n
    with pytest.raises(pg.PgnError):
        pg.csr_linear_solve(ctx, rp, bad, v.copy(), np.ones(n))
    with pytest.raises(pg.PgnError):  # decreasing row pointers
        rp2 = rp.copy()
        rp2[2] = rp2[1] - 1
        pg.csr_linear_solve(ctx, rp2, ci, v.copy(), np.ones(n))
    with pytest.raises(pg.PgnError):  # no diagonal in row 0
        pg.csr_linear_solve(ctx, np.array([0, 0, 1], np.int64), np.array([1], np.int32),
                            np.array([1.0]), np.ones(2), layout=pg.LAYOUT_FULL)
    with pytest.raises(pg.PgnError):  # block Jacobi is not offered here
        pg.csr_linear_solve(ctx, rp, ci, v.copy(), np.ones(n), precond=1)


def test_csr_solve_takes_the_models_exported_hessian(ctx):
    """The drop-in path: a model's Hessian exported in AMatrixSd's layout and
    solved by the caller-matrix entry point matches the model's own FP64
    Jacobi PCG on the same system."""
    _say("csr solve of an exported model Hessian")
    s = scenes.cantilever(dims=(4, 4, 16), scale=(0.1, 0.1, 0.4))
    G = pgn_model(ctx, s, pg.FP64)
    for _ in range(s.stages):
        G.step_load()
    G.set_x(perturbed_x(s, 1e-4))
    rp, ci, v = G.hessian(pg.LAYOUT_LOWER)
    b = -G.gradient()
    x, r = pg.csr_linear_solve(ctx, rp, ci, v.copy(), b, tol=1e-8, max_iters=5000)
    xm, rm = G.linear_solve(b, tol=1e-8, max_iters=5000, precond=0)
    _say(f"  csr {r}  model {rm}")
    assert r["status"] == rm["status"] == 0
    assert abs(r["iterations"] - rm["iterations"]) <= 2
    assert np.linalg.norm(x - xm) <= 1e-6 * np.linalg.norm(xm)
    G.close()
