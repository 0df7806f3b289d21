"""The oracle's EigenLDLT restatement (LinearSolver_EigenLDLT.cpp:49-153 with
LinearSolver::GetMatrixData, LinearSolver.cpp:57-66), pinned by what the
mathematics fixes:

- SPD systems: the solution numpy.linalg.solve gives, first try;
- an indefinite diagonal matrix: the regularization try at which the
  cumulative shift regT (1 + 10 + ... + 10^(i-1)) first makes every pivot
  positive, and the solution of the shifted system (closed form);
- a pivot below sinT = max(min(1e-6 |A|, 1e-6), 1e-9): no try succeeds;
- the zero matrix: FAILURE, x keeps its input;
- the lower-CSR entry point on a tet model's Hessian: normA over every
  stored entry (AMatrixSd keeps whole diagonal blocks), the same solution as
  the dense solve of the full matrix;
- Newton with EigenLDLT against Newton with Jacobi CG at a tight linear
  tolerance: the same equilibrium.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import orc
from tests.helpers import oracle_model, small_beam


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    Q = rng.standard_normal((n, n))
    return Q @ Q.T + n * np.eye(n)


@pytest.mark.parametrize("n", [1, 5, 40])
def test_spd_matches_numpy(n):
    A = _spd(n, n)
    b = np.random.default_rng(1).standard_normal(n)
    x = np.zeros(n)
    st, reg = orc.dense_ldlt_solve(A, b, x, 1e-9, 10)
    assert st == 0 and reg == 0
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-12, atol=1e-14)


def test_indefinite_diagonal_regularization_count():
    d = np.array([-1.0, 1.0, 1.0, 1.0])
    A = np.diag(d)
    b = np.array([1.0, 0.5, 0.25, 0.125])
    nA = np.linalg.norm(d)
    regT = 1e-6 * nA
    # first try i whose cumulative shift regT (10^i - 1) / 9 exceeds 1
    i = next(k for k in range(1, 11) if regT * (10.0 ** k - 1) / 9 > 1.0)
    shift = regT * (10.0 ** i - 1) / 9
    x = np.zeros(4)
    st, reg = orc.dense_ldlt_solve(A, b, x, 1e-9, 10)
    assert st == 0 and reg == i == 7
    assert np.allclose(x, b / (d + shift), rtol=1e-12)


def test_pivot_below_singular_threshold_fails():
    A = np.array([[1e-12]])
    x = np.array([3.0])
    st, reg = orc.dense_ldlt_solve(A, np.array([1.0]), x, 1e-9, 10)
    assert st == 1 and reg == 9
    assert x[0] == 3.0  # no factorization succeeded: x keeps its input


def test_zero_matrix_fails_and_keeps_x():
    x = np.array([1.0, 2.0, 3.0])
    st, reg = orc.dense_ldlt_solve(np.zeros((3, 3)), np.ones(3), x, 1e-9, 10)
    assert st == 1
    assert np.array_equal(x, [1.0, 2.0, 3.0])


def test_sparse_entry_uses_stored_norm_and_matches_dense():
    s = small_beam(dims=(2, 2, 3))
    O = oracle_model(s)
    while O.step_load():
        pass
    rp, ci, v = O.hessian_lower()
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    L = sp.csr_matrix((np.where(ci <= rows, v, 0.0), (rows, ci)), shape=(n, n)).toarray()
    Afull = L + np.tril(L, -1).T
    b = -O.gradient()
    x = np.zeros(n)
    st, reg = orc.sparse_ldlt_solve(rp, ci, v, b, x, 1e-9, 10)
    assert st == 0 and reg == 0
    assert np.allclose(x, np.linalg.solve(Afull, b), rtol=1e-10, atol=1e-14)
    # whole diagonal blocks are stored: the norm counts their upper entries
    assert np.any(ci > rows)
    assert np.sqrt(np.sum(v * v)) > np.sqrt(np.sum(L * L))


def test_newton_ldlt_matches_newton_cg():
    s = small_beam(dims=(2, 2, 5))
    opts = dict(tolMaxError=1e-9, lsTolMaxError=1e-12, lsNumMaxIters=5000, tolMinStep=0.0,
                tolMinImpr=0.0, ls_noise_guard=1e-12)
    A = oracle_model(s)
    ra = A.newton(**opts, ls_solver=2)
    B = oracle_model(s)
    rb = B.newton(**opts, ls_solver=1)
    assert ra["result"] == "SUCCESS" and rb["result"] == "SUCCESS", (ra, rb)
    assert ra["cg_iterations"] == 0 and rb["cg_iterations"] > 0
    assert np.abs(A.get_x() - B.get_x()).max() <= 1e-8 * s.bbox_diag()
