"""QPSolverType::LBFGS and ::Steepest on the GPU single-model driver against
the oracle (ComputeStep_LBFGS, OptimSolver.cpp:522-591, corrected as
documented in include/pgn.h; steepest descent dx = -g, :332-334): no Hessian
and no linear solve, the converged positions within 1e-6 x the bounding-box
diagonal of the oracle's Newton equilibrium.  The first LBFGS iterations
follow the oracle's own LBFGS run (same directions up to FP64 rounding of
the dot products)."""
import sys

import numpy as np
import pytest

import physim_amd as pg
from physim_amd import scenes
from tests.helpers import oracle_model, pgn_model

pytestmark = pytest.mark.gpu

OPTS = dict(tolMaxError=1e-7, tolMinImpr=0.0, tolMinStep=0.0, ls_noise_guard=1e-12,
            numMaxIters=20000)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("material", [0, 1])
def test_lbfgs_matches_oracle(ctx, material):
    _say(f"lbfgs material={material}")
    s = scenes.cantilever(dims=(2, 2, 4), scale=(0.1, 0.1, 0.2), material=material)
    O = oracle_model(s)
    ro = O.newton(**OPTS)  # Newton equilibrium
    OL = oracle_model(s)
    rol = OL.newton(**OPTS, qpSolverType=pg.QP_LBFGS)
    G = pgn_model(ctx, s, pg.FP64)
    r = G.newton(**OPTS, qpSolverType=pg.QP_LBFGS)
    _say(f"  gpu {r['result']} {r['iterations']} its; oracle lbfgs {rol['result']} "
         f"{rol['iterations']} its")
    assert ro["result"] == rol["result"] == r["result"] == "SUCCESS"
    assert r["pcg_iterations"] == 0 and r["hessian_evals"] == 0
    assert abs(r["iterations"] - rol["iterations"]) <= 0.25 * rol["iterations"]
    assert np.abs(G.get_x() - O.get_x()).max() <= 1e-6 * s.bbox_diag()
    # a short run follows the oracle's LBFGS iterates
    G.reset()
    OL2 = oracle_model(s)
    G.newton(**dict(OPTS, numMaxIters=12), qpSolverType=pg.QP_LBFGS)
    OL2.newton(**dict(OPTS, numMaxIters=12), qpSolverType=pg.QP_LBFGS)
    assert np.abs(G.get_x() - OL2.get_x()).max() <= 1e-9 * s.bbox_diag()
    G.close()


def test_steepest_matches_oracle(ctx):
    _say("steepest descent")
    s = scenes.cantilever(dims=(1, 1, 2), scale=(0.1, 0.1, 0.2))
    opts = dict(OPTS, tolMaxError=1e-6, numMaxIters=50000)
    O = oracle_model(s)
    ro = O.newton(**opts)
    G = pgn_model(ctx, s, pg.FP64)
    r = G.newton(**opts, qpSolverType=pg.QP_STEEPEST)
    _say(f"  gpu {r['result']} {r['iterations']} its")
    assert r["result"] == "SUCCESS" and r["hessian_evals"] == 0
    assert np.abs(G.get_x() - O.get_x()).max() <= 1e-6 * s.bbox_diag()
    G.close()


def test_batches_take_newton_steps_only(ctx):
    B = pg.Batch(ctx, pg.FP64)
    B.add_scene(scenes.cantilever(dims=(1, 1, 2), scale=(0.1, 0.1, 0.2)))
    B.build()
    with pytest.raises(pg.PgnError, match="PGN_QP_NEWTON"):
        B.newton(**OPTS, qpSolverType=pg.QP_LBFGS)
    B.close()
