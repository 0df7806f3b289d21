"""The 1.28M-tet beam (40x40x160, lognormal E, 5% jitter) at full size on the
GPU, through properties that do not depend on the problem size:

- element E, g, H of a random sample of 2,000 elements against the oracle's
  single-element routine (EnergyElement_FEM.cpp:125-222), at a smooth large
  deformation: FP64 kernel relative 1e-10, FP32 kernel relative 1e-5
  (normwise per element);
- translation invariance of the assembled gradient and Hessian of the
  unclamped beam (internal forces sum to zero; H annihilates rigid
  translations), and exact symmetry of the stored BSR3 blocks;
- the mixed-precision linear solve's residual recomputed in FP64 on the host
  from the exported matrix;
- the mixed-precision Newton solve against the plain FP64 path, within
  1e-6 x bounding-box diagonal.
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import physim_amd as pg
from oracle import orc
from physim_amd import scenes
from tests.helpers import pgn_model

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


@pytest.fixture(scope="module")
def beam():
    return scenes.config2()


def _bent(s):
    """A smooth large deformation: the beam bent down and twisted a little,
    plus small noise; every element stays positively oriented."""
    X = s.X0
    z = X[:, 2] / X[:, 2].max()
    u = np.zeros_like(X)
    u[:, 1] = -0.08 * z ** 2
    ang = 0.3 * z
    c, sn = np.cos(ang), np.sin(ang)
    xc, yc = X[:, 0] - 0.05, X[:, 1] - 0.05
    u[:, 0] = (c - 1) * xc - sn * yc
    u[:, 1] += sn * xc + (c - 1) * yc
    u += 1e-4 * np.random.default_rng(3).standard_normal(X.shape)
    u[s.fixed.astype(bool)] = 0.0
    return X + u


@pytest.mark.parametrize("precision,tol", [(pg.FP64, 1e-10), (pg.MIXED, 1e-5)])
def test_sampled_elements_match_oracle(ctx, beam, precision, tol):
    _say(f"full size: sampled elements precision={precision}")
    s = beam
    x = _bent(s)
    G = pgn_model(ctx, s, pg.FP64)
    G.set_x(x)
    E, g, H = G.element_egh(precision)
    rng = np.random.default_rng(11)
    for e in rng.choice(s.m, size=2000, replace=False):
        nodes = s.T[e]
        D, V0 = orc.tet_rest(s.X0[nodes])
        Eo, go, Ho, inv = orc.tet_egh(x[nodes], D, V0, int(s.material[e]), s.lame1[e],
                                      s.lame2[e])
        assert not inv, e
        assert abs(E[e] - Eo) <= tol * abs(Eo) + 1e-300, e
        assert _rel(g[e], go) <= tol, e
        assert _rel(H[e], Ho) <= tol, e
    G.close()


@pytest.mark.parametrize("precision,tol", [(pg.FP64, 1e-11), (pg.MIXED, 1e-4)])
def test_translation_invariance_and_symmetry(ctx, beam, precision, tol):
    _say(f"full size: invariance and symmetry precision={precision}")
    s = beam
    free = scenes.Scene("beam_free", s.X0, s.T, s.material, s.lame1, s.lame2, s.density,
                        np.zeros_like(s.fixed), (0.0, 0.0, 0.0), 1)
    G = pgn_model(ctx, free, precision)
    G.set_x(_bent(s))
    g = G.gradient().reshape(-1, 3)
    assert np.abs(g.sum(0)).max() <= 1e-10 * np.abs(g).sum()  # FP64 in both paths
    rp, ci, v = G.hessian(pg.LAYOUT_BSR3)
    B = v.reshape(-1, 3, 3)
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    # exact block symmetry: block (i, j) is the transpose of block (j, i)
    order = np.lexsort((rows, ci))  # blocks sorted by (column, row) = transposed order
    assert np.array_equal(ci[order], rows) and np.array_equal(rows[order], ci)
    assert np.array_equal(B[order], np.transpose(B, (0, 2, 1)))
    A = sp.bsr_matrix((B, ci, rp), shape=(G.N, G.N)).tocsr()
    scale = np.abs(A).sum(1).max()
    for k in range(3):
        t = np.zeros(G.N)
        t[k::3] = 1.0
        assert np.abs(A @ t).max() <= tol * scale, k
    G.close()


def test_mixed_linear_solve_residual(ctx, beam):
    _say("full size: mixed linear solve residual")
    s = beam
    G = pgn_model(ctx, s, pg.MIXED)
    while G.step_load():
        pass
    b = -G.gradient()  # full gravity at rest: the SPD linear-elastic stiffness
    rp, ci, v = G.hessian(pg.LAYOUT_BSR3)
    x, r = G.linear_solve(b, None, tol=1e-6, max_iters=20000, reg_iters=1, precond=1)
    assert r["status"] == 0, r
    A = sp.bsr_matrix((v.reshape(-1, 3, 3), ci, rp), shape=(G.N, G.N))
    res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    _say(f"  {r}, host residual {res:.3e}")
    assert res <= 1e-6 * 1.001
    G.close()


def test_newton_mixed_matches_fp64(ctx, beam):
    _say("full size: Newton mixed vs FP64")
    s = beam
    opts = dict(pg.CONVERGE_OPTS, precond=1, adaptive_tol=1, lsNumMaxIters=2000)
    M = pgn_model(ctx, s, pg.MIXED)
    rm = M.newton(**opts)
    F = pgn_model(ctx, s, pg.FP64)
    rf = F.newton(**dict(pg.CONVERGE_OPTS, precond=0))
    _say(f"  mixed {rm}\n  fp64 {rf}")
    assert rm["result"] == "SUCCESS" and rf["result"] == "SUCCESS"
    assert np.abs(M.get_x() - F.get_x()).max() <= 1e-6 * s.bbox_diag()
    M.close()
    F.close()
