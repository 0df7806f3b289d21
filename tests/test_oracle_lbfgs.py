"""LBFGS (QPSolverType::LBFGS, ComputeStep_LBFGS, OptimSolver.cpp:522-591) in
the oracle, corrected as documented in oracle.h (previous x and g updated at
every step, second loop modulo the history length, pair count saturating at
it).

The two-loop recursion is pinned by the dense BFGS inverse update it is
equivalent to (Nocedal & Wright, Numerical Optimization, 2nd ed., eq. 6.17
and Alg. 7.4): H0 = gamma I with gamma = s.y / y.y of the newest pair, then
for each pair from oldest to newest H <- (I - rho s y^T) H (I - rho y s^T) +
rho s s^T, rho = 1 / (y.s); the direction is -H g.  A reversed loop order, a
wrong ring index or a dropped term changes the product."""
import numpy as np
import pytest

from oracle import orc
from physim_amd import scenes
from tests.helpers import oracle_model


def _dense(S, Y, order, g):
    s_new, y_new = S[order[-1]], Y[order[-1]]
    n = g.size
    H = (s_new @ y_new) / (y_new @ y_new) * np.eye(n)
    for i in order:
        s, y = S[i], Y[i]
        rho = 1.0 / (y @ s)
        V = np.eye(n) - rho * np.outer(y, s)
        H = V.T @ H @ V + rho * np.outer(s, s)
    return -H @ g


@pytest.mark.parametrize("count,newest", [(1, 0), (2, 1), (3, 2), (3, 0), (3, 1), (2, 0)])
def test_two_loop_matches_dense_bfgs_update(count, newest):
    rng = np.random.default_rng(10 * count + newest)
    n, M = 9, 3
    B = rng.standard_normal((n, n))
    A = B @ B.T + n * np.eye(n)  # SPD: y = A s gives y.s > 0
    S = rng.standard_normal((M, n))
    Y = S @ A.T
    g = rng.standard_normal(n)
    order = [(newest - count + 1 + i) % M for i in range(count)]  # oldest .. newest
    dx = orc.lbfgs_direction(S, Y, g, count, newest)
    want = _dense(S, Y, order, g)
    np.testing.assert_allclose(dx, want, rtol=1e-11, atol=1e-13 * np.abs(want).max())
    assert dx @ g < 0  # a descent direction


def test_three_pairs_on_a_quadratic_reproduce_the_inverse_on_their_span():
    """Secant equations: with the newest pair, H y_new = s_new exactly."""
    rng = np.random.default_rng(3)
    n = 6
    B = rng.standard_normal((n, n))
    A = B @ B.T + np.eye(n)
    S = rng.standard_normal((3, n))
    Y = S @ A.T
    dx = orc.lbfgs_direction(S, Y, -Y[2], 3, 2)  # -H (-y_new) = H y_new
    np.testing.assert_allclose(dx, S[2], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("material", [0, 1])
def test_lbfgs_reaches_newtons_equilibrium(material):
    s = scenes.cantilever(dims=(2, 2, 4), scale=(0.1, 0.1, 0.2), material=material)
    opts = dict(tolMaxError=1e-7, tolMinImpr=0.0, tolMinStep=0.0, ls_noise_guard=1e-12,
                numMaxIters=20000)
    N = oracle_model(s)
    rn = N.newton(**opts)
    L = oracle_model(s)
    rl = L.newton(**opts, qpSolverType=2)
    assert rn["result"] == "SUCCESS", rn
    assert rl["result"] == "SUCCESS", rl
    assert rl["iterations"] > rn["iterations"]  # first order: many more steps
    assert rl["cg_iterations"] == 0  # no linear solves
    assert np.abs(L.get_x() - N.get_x()).max() <= 1e-6 * s.bbox_diag()
