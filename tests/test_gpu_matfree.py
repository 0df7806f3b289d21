"""The matrix-free Hessian product (element products over the stored element
Hessians, gathered per node in element order) against the assembled BSR3
product and the oracle's dense Hessian (the product of each CG iteration,
LinearSolver_EigenCG.cpp:89).  Tolerances: FP64 models 1e-12 relative
(normwise; the two sum the same terms in different orders), mixed models
1e-5 (FP32 values).  Under PGN_MATFREE=1 the PCG uses it: Newton solves of
the config-1 beam reach the oracle's equilibrium within 1e-6 x bbox, and a
regularized solve keeps the shift in the product."""
import sys

import numpy as np
import pytest

import physim_amd as pg
from oracle import orc
from physim_amd import scenes
from tests.helpers import oracle_model, perturbed_x, pgn_model

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("precision,tol", [(pg.FP64, 1e-12), (pg.MIXED, 1e-5)])
def test_matrix_free_product_matches_assembled_and_oracle(ctx, precision, tol):
    _say(f"matrix-free product precision={precision}")
    s = scenes.cantilever(dims=(3, 3, 8), scale=(0.1, 0.1, 0.25), young_sigma=0.3,
                          jitter=0.05, seed=4)
    s.material[::3] = 1
    O = oracle_model(s)
    G = pgn_model(ctx, s, precision)
    for _ in range(s.stages):
        O.step_load()
        G.step_load()
    x = perturbed_x(s, 1e-3)
    O.set_x(x)
    G.set_x(x)
    A = orc.lower_to_dense(*O.hessian_lower())
    rng = np.random.default_rng(6)
    for _ in range(3):
        p = rng.standard_normal(G.N)
        ya = G.hessian_product(p)
        ym = G.hessian_product(p, matrix_free=True)
        yo = A @ p
        _say(f"  assembled {_rel(ya, yo):.2e}  matrix-free {_rel(ym, yo):.2e}  "
             f"apart {_rel(ym, ya):.2e}")
        assert _rel(ya, yo) <= tol and _rel(ym, yo) <= tol and _rel(ym, ya) <= tol
    G.close()


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
def test_matrix_free_newton_config1(ctx, precision, monkeypatch):
    _say(f"matrix-free newton config1 precision={precision}")
    s = scenes.config1()
    O = oracle_model(s)
    ro = O.newton(**pg.CONVERGE_OPTS)
    monkeypatch.setenv("PGN_MATFREE", "1")  # read when the model is created
    G = pgn_model(ctx, s, precision)
    r = G.newton(**pg.CONVERGE_OPTS)
    _say(f"  gpu {r}  oracle {ro['result']} {ro['iterations']}")
    assert ro["result"] == "SUCCESS"
    assert r["result"] == "SUCCESS" and r["optimality"] <= 1e-8
    assert np.abs(G.get_x() - O.get_x()).max() <= 1e-6 * s.bbox_diag()
    G.close()


def test_matrix_free_regularized_solve(ctx, monkeypatch):
    """A squashed state whose Hessian has a negative eigenvalue lambda with
    eigenvector v: with b = diag(A) v and one Jacobi-PCG iteration per try,
    the first direction is p = v, so alpha = r.z / p.Ap < 0 and b.x < 0 until
    the cumulative shift regT (10^i - 1) / 9 passes |lambda|.  The
    matrix-free PCG carries the shift in its product: it takes the same
    try and the same step as the assembled one."""
    _say("matrix-free regularized solve")
    s = scenes.cantilever(dims=(2, 2, 6), scale=(0.1, 0.1, 0.3))
    x = s.X0.copy()
    x[:, 2] *= 0.55  # squashed along the beam: indefinite Hessian
    O = oracle_model(s)
    for _ in range(s.stages):
        O.step_load()
    O.set_x(x)
    A = orc.lower_to_dense(*O.hessian_lower())
    w, V = np.linalg.eigh(A)
    assert w[0] < 0
    b = np.diag(A) * V[:, 0]
    out = []
    for mf in ("0", "1"):
        monkeypatch.setenv("PGN_MATFREE", mf)
        monkeypatch.setenv("PGN_PCG", "split")
        G = pgn_model(ctx, s, pg.FP64)
        for _ in range(s.stages):
            G.step_load()
        G.set_x(x)
        xs, r = G.linear_solve(b, tol=1e-10, max_iters=1, precond=0)
        out.append((xs, r))
        G.close()
    (xa, ra), (xm, rm) = out
    _say(f"  assembled {ra}  matrix-free {rm}")
    assert ra["status"] == rm["status"] and ra["reg_used"] == rm["reg_used"] > 0
    assert np.linalg.norm(xm - xa) <= 1e-10 * np.linalg.norm(xa)
    assert b @ xm >= 0
