"""libpgn tet path on the GPU vs the FP64 oracle, through the C ABI.

Tolerances: FP64 kernels relative 1e-10 (normwise per element) against the
oracle; the all-FP32 element kernel relative 1e-5 (the north-star element
E/g/H bound); assembled quantities relative 1e-12 (FP64) / 1e-5 (FP32
Hessian); converged displacement within 1e-6 x bounding-box diagonal.
"""
import os
import sys

import numpy as np
import pytest
import scipy.sparse

import physim_amd as pg
from oracle import orc
from physim_amd import scenes
from tests.helpers import oracle_model, perturbed_x, pgn_model, small_beam

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "material_golden.npz")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


@pytest.mark.parametrize("precision,tol", [(pg.FP64, 1e-10), (pg.MIXED, 1e-5)])
@pytest.mark.parametrize("amp", [1e-4, 3e-3])
def test_element_egh_matches_oracle(ctx, precision, tol, amp):
    _say(f"element_egh precision={precision} amp={amp}")
    s = small_beam(dims=(3, 3, 6), mixed_materials=True, seed=11)
    x = perturbed_x(s, amp, seed=12)
    O = oracle_model(s)
    O.set_x(x)
    Eo, go, Ho = O.element_egh()
    G = pgn_model(ctx, s, pg.FP64)
    G.set_x(x)
    E, g, H = G.element_egh(precision)
    for e in range(s.m):
        assert abs(E[e] - Eo[e]) <= tol * abs(Eo[e]) + 1e-300, e
        assert _rel(g[e], go[e]) <= tol, e
        assert _rel(H[e], Ho[e]) <= tol, e
        assert np.array_equal(H[e], H[e].T)


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
def test_element_from_golden_material(ctx, precision):
    """Unit rest tet (Dm = I, V0 = 1/6) deformed by the golden F: E, g, H are
    the reference Maple Psi, P, dP/dF pushed through DFDx."""
    _say(f"golden material precision={precision}")
    gd = np.load(GOLD)
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
    D, V0 = orc.tet_rest(X)
    idx = [i for i in range(len(gd["psi"]))
           if np.linalg.det(gd["Fv"][i].reshape(3, 3).T) > 1e-2]
    m = len(idx)
    X0 = np.tile(X, (m, 1))
    T = (np.arange(4 * m, dtype=np.int32)).reshape(m, 4)
    M = pg.TetModel(ctx, X0, T, gd["kind"][idx], gd["lame1"][idx], gd["lame2"][idx],
                    1.0, pg.FP64)
    xs = []
    for i in idx:
        F = gd["Fv"][i].reshape(3, 3).T
        xs.append(np.vstack([np.zeros(3), F[:, 0], F[:, 1], F[:, 2]]))
    M.set_x(np.vstack(xs))
    E, g, H = M.element_egh(precision)
    tol = 1e-9 if precision == pg.FP64 else 1e-5
    for k, i in enumerate(idx):
        Cg = gd["C"][i].reshape(9, 9)
        assert abs(E[k] - gd["psi"][i] / 6) <= tol * abs(gd["psi"][i] / 6) + 1e-12
        assert _rel(g[k], D.T @ gd["P"][i] / 6) <= tol
        assert _rel(H[k], D.T @ Cg @ D / 6) <= tol


def test_inverted_element_energy_is_inf(ctx):
    _say("inverted element")
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
    M = pg.TetModel(ctx, X, np.array([[0, 1, 2, 3]], np.int32), 0, 1.0, 1.0, 1.0)
    x = X.copy()
    x[3, 2] = -0.5
    M.set_x(x)
    assert np.isinf(M.energy())
    for prec in (pg.FP64, pg.MIXED):
        E, g, H = M.element_egh(prec)
        assert np.isinf(E[0]) and np.isfinite(g).all() and np.isfinite(H).all()


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
def test_energy_gradient_match_oracle(ctx, precision):
    _say(f"energy/gradient precision={precision}")
    s = small_beam(dims=(3, 4, 8), mixed_materials=True, seed=21)
    rng = np.random.default_rng(22)
    s.forces = 0.01 * rng.standard_normal((s.n, 3))
    O = oracle_model(s)
    G = pgn_model(ctx, s, precision)
    for _ in range(4):
        O.step_load()
        G.step_load()
    x = perturbed_x(s, 2e-3, seed=23)
    O.set_x(x)
    G.set_x(x)
    Eo, go = O.energy(), O.gradient()
    E, g = G.energy(), G.gradient()
    assert abs(E - Eo) <= 1e-12 * (abs(Eo) + 1.0)
    assert np.abs(g - go).max() <= 1e-11 * np.abs(go).max()
    assert np.all(g[s.fixed.reshape(-1).astype(bool)] == 0.0)
    # bitwise deterministic
    assert np.array_equal(g, G.gradient()) and G.energy() == E


@pytest.mark.parametrize("precision,tol", [(pg.FP64, 1e-12), (pg.MIXED, 1e-5)])
def test_hessian_matches_oracle(ctx, precision, tol):
    _say(f"hessian precision={precision}")
    s = small_beam(dims=(3, 3, 7), mixed_materials=True, seed=31)
    x = perturbed_x(s, 2e-3, seed=32)
    O = oracle_model(s)
    G = pgn_model(ctx, s, precision)
    O.set_x(x)
    G.set_x(x)
    rpo, cio, vo = O.hessian_lower()
    rp, ci, v = G.hessian(pg.LAYOUT_LOWER)
    assert np.array_equal(rp, rpo) and np.array_equal(ci, cio)
    assert _rel(v, vo) <= tol
    assert np.array_equal(v, G.hessian(pg.LAYOUT_LOWER)[2])  # deterministic
    # full layout is the symmetric completion
    rpf, cif, vf = G.hessian(pg.LAYOUT_FULL)
    A = orc.lower_to_dense(rp, ci, v)
    Af = np.zeros_like(A)
    for r in range(A.shape[0]):
        Af[r, cif[rpf[r]:rpf[r + 1]]] = vf[rpf[r]:rpf[r + 1]]
    assert np.array_equal(Af, Af.T)
    assert _rel(Af, A) <= tol


def test_linear_solve_fp64_matches_oracle_cg(ctx):
    _say("linear solve fp64")
    s = small_beam(dims=(3, 3, 8), seed=41)
    x = perturbed_x(s, 1e-3, seed=42)
    O = oracle_model(s)
    G = pgn_model(ctx, s, pg.FP64)
    O.set_x(x)
    G.set_x(x)
    rp, ci, v = O.hessian_lower()
    b = np.random.default_rng(43).standard_normal(s.n * 3)
    b[s.fixed.reshape(-1).astype(bool)] = 0.0
    xo, ito, erro = orc.cg(rp, ci, v, b, np.zeros_like(b), 1e-9, 1000)
    xg, info = G.linear_solve(b, tol=1e-9, max_iters=1000)
    assert info["status"] == 0
    assert abs(info["iterations"] - ito) <= 2
    assert np.linalg.norm(xg - xo) <= 1e-7 * np.linalg.norm(xo)
    A = orc.lower_to_dense(rp, ci, v)
    assert np.linalg.norm(A @ xg - b) <= 1.5e-9 * np.linalg.norm(b)
    xg2, _ = G.linear_solve(b, tol=1e-9, max_iters=1000)
    assert np.array_equal(xg, xg2)


@pytest.mark.parametrize("precond", [0, 1])
def test_linear_solve_mixed_reaches_fp64_residual(ctx, precond):
    _say(f"linear solve mixed precond={precond}")
    s = small_beam(dims=(3, 3, 8), seed=51)
    x = perturbed_x(s, 1e-3, seed=52)
    O = oracle_model(s)
    G = pgn_model(ctx, s, pg.MIXED)
    O.set_x(x)
    G.set_x(x)
    b = np.random.default_rng(53).standard_normal(s.n * 3)
    b[s.fixed.reshape(-1).astype(bool)] = 0.0
    xg, info = G.linear_solve(b, tol=1e-9, max_iters=3000, precond=precond)
    assert info["status"] == 0 and info["refinements"] >= 1
    # residual against the FP32-assembled matrix, in FP64
    rp, ci, v = G.hessian(pg.LAYOUT_LOWER)
    A = orc.lower_to_dense(rp, ci, v)
    assert np.linalg.norm(A @ xg - b) <= 1e-9 * np.linalg.norm(b) * 1.01


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
def test_sell_and_bsr_spmv_solve_the_same_system(ctx, precision, monkeypatch):
    """The PCG over the SELL-64 copy of the matrix (default) and over BSR3
    (PGN_SELL=0, read at model creation): both reach the FP64 residual of
    the same assembled matrix (so their solutions differ by at most the
    conditioning times 2e-10; checked loosely at 1e-4).  400 nodes (5 x 5
    per section) fill six slices of 64 rows and a ragged seventh, with rows
    of 4 to 19 blocks."""
    _say(f"sell vs bsr precision={precision}")
    s = small_beam(dims=(4, 4, 15), seed=61)
    x = perturbed_x(s, 1e-3, seed=62)
    b = np.random.default_rng(63).standard_normal(s.n * 3)
    b[s.fixed.reshape(-1).astype(bool)] = 0.0
    sols = []
    for sell in ("1", "0"):
        monkeypatch.setenv("PGN_SELL", sell)
        G = pgn_model(ctx, s, precision)
        G.set_x(x)
        xg, info = G.linear_solve(b, tol=1e-10, max_iters=5000, precond=1)
        assert info["status"] == 0, info
        rp, ci, v = G.hessian(pg.LAYOUT_LOWER)
        A = orc.lower_to_dense(rp, ci, v)
        assert np.linalg.norm(A @ xg - b) <= 1e-10 * np.linalg.norm(b) * 1.01
        sols.append(xg)
        G.close()
    assert np.linalg.norm(sols[0] - sols[1]) <= 1e-4 * np.linalg.norm(sols[1])


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
def test_sell_16bit_columns_match_32bit(ctx, precision, monkeypatch):
    """The SELL-64 copy with block columns as 16-bit offsets from the row
    (default where they fit) and as 32-bit columns (PGN_SELL_C16=0, read when
    the copy is first built): the same products in the same order, so the
    same iterations and bit-identical solutions.  The 400-node beam's ragged
    seventh slice has 48 lanes past the last row, whose padding offsets the
    first-block loads read."""
    _say(f"sell 16-bit columns precision={precision}")
    s = small_beam(dims=(4, 4, 15), seed=61)
    x = perturbed_x(s, 1e-3, seed=62)
    b = np.random.default_rng(64).standard_normal(s.n * 3)
    b[s.fixed.reshape(-1).astype(bool)] = 0.0
    out = []
    for c16, width in (("1", 2), ("0", 4)):
        monkeypatch.setenv("PGN_SELL_C16", c16)
        G = pgn_model(ctx, s, precision)
        G.set_x(x)
        xg, info = G.linear_solve(b, tol=1e-10, max_iters=5000, precond=1)
        assert info["status"] == 0, info
        assert G.profile(5)[0] == width
        out.append((xg, info["iterations"]))
        G.close()
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])


@pytest.mark.parametrize("dims", [None, (1, 1, 2)])
def test_sell_models_smaller_than_a_slice(ctx, dims):
    """Models of 4 and 12 nodes: one SELL-64 slice with 60 and 52 lanes past
    the last row.  The PCG solve reaches the FP64 residual of the exported
    matrix, and equals numpy's dense solve of it within its conditioning."""
    _say(f"sell tiny model dims={dims}")
    if dims is None:
        X0 = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.1, 0.0], [0.02, 0.03, 0.1]])
        l1, l2 = orc.lame(1e5, 0.3)
        fixed = np.zeros((4, 3), dtype=np.uint8)
        fixed[:3] = 1
        s = scenes.Scene("one_tet", X0, np.array([[0, 1, 2, 3]], dtype=np.int32),
                         np.array([0], dtype=np.uint8), np.array([l1]), np.array([l2]),
                         np.array([1000.0]), fixed, scenes.G_EARTH, 1)
    else:
        s = small_beam(dims=dims, seed=68)
    assert s.n < 64
    for precision in (pg.FP64, pg.MIXED):
        G = pgn_model(ctx, s, precision)
        G.set_x(perturbed_x(s, 1e-3, seed=69))
        b = np.random.default_rng(70).standard_normal(s.n * 3)
        b[s.fixed.reshape(-1).astype(bool)] = 0.0
        xg, info = G.linear_solve(b, tol=1e-10, max_iters=1000, precond=1)
        assert info["status"] == 0, info
        assert G.profile(5)[0] == 2
        rp, ci, v = G.hessian(pg.LAYOUT_LOWER)
        A = orc.lower_to_dense(rp, ci, v)
        assert np.linalg.norm(A @ xg - b) <= 1e-10 * np.linalg.norm(b) * 1.01
        xd = np.linalg.solve(A, b)
        assert np.linalg.norm(xg - xd) <= 1e-10 * np.linalg.cond(A) * np.linalg.norm(xd) * 1.01
        G.close()


def test_sell_wide_band_keeps_32bit_columns(ctx):
    """A 36,421-node beam with its node ids shuffled: blocks lie up to ~36,000
    nodes off the diagonal, so the SELL-64 copy keeps 32-bit columns by
    itself, and the solve reaches the FP64 residual of the assembled matrix
    (checked with a sparse product of the exported BSR3 matrix)."""
    _say("sell wide band")
    s = small_beam(dims=(10, 10, 300), seed=65)
    perm = np.random.default_rng(66).permutation(s.n)
    X0 = np.empty_like(s.X0)
    X0[perm] = s.X0
    fixed = np.empty_like(s.fixed)
    fixed[perm] = s.fixed
    sp_ = scenes.Scene("shuffled beam", X0, perm[s.T].astype(np.int32), s.material, s.lame1,
                       s.lame2, s.density, fixed, s.gravity, s.stages)
    G = pgn_model(ctx, sp_, pg.MIXED)
    b = np.random.default_rng(67).standard_normal(s.n * 3)
    b[fixed.reshape(-1).astype(bool)] = 0.0
    xg, info = G.linear_solve(b, tol=1e-8, max_iters=20000, precond=1)
    _say(f"  {info}")
    assert info["status"] == 0, info
    assert G.profile(5)[0] == 4
    rp, ci, v = G.hessian(pg.LAYOUT_BSR3)
    A = scipy.sparse.bsr_matrix((v.reshape(-1, 3, 3), ci, rp), shape=(G.N, G.N))
    assert np.linalg.norm(A @ xg - b) <= 1e-8 * np.linalg.norm(b) * 1.01
    G.close()


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
def test_newton_small_beam_matches_oracle(ctx, precision):
    _say(f"newton small beam precision={precision}")
    s = scenes.cantilever(dims=(4, 4, 16), scale=(0.1, 0.1, 0.25), young_sigma=0.3,
                          jitter=0.05, seed=0)
    O = oracle_model(s)
    ro = O.newton(**pg.CONVERGE_OPTS)
    assert ro["result"] == "SUCCESS"
    G = pgn_model(ctx, s, precision)
    r = G.newton(**pg.CONVERGE_OPTS)
    _say(f"  gpu {r}  oracle {ro}")
    assert r["result"] == "SUCCESS" and r["optimality"] <= 1e-8
    dx = np.abs(G.get_x() - O.get_x()).max()
    assert dx <= 1e-6 * s.bbox_diag()


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
def test_newton_config1_beam(ctx, precision):
    """Config 1: 10x10x40 CoNH beam, 20,000 tets, gravity over 10 stages,
    residual 1e-8; displacement within 1e-6 bbox diagonal of the oracle."""
    _say(f"newton config1 precision={precision}")
    s = scenes.config1()
    O = oracle_model(s)
    ro = O.newton(**pg.CONVERGE_OPTS)
    G = pgn_model(ctx, s, precision)
    r = G.newton(**pg.CONVERGE_OPTS)
    _say(f"  gpu {r}  oracle {ro}")
    assert ro["result"] == "SUCCESS"
    assert r["result"] == "SUCCESS" and r["optimality"] <= 1e-8
    assert np.abs(G.get_x() - O.get_x()).max() <= 1e-6 * s.bbox_diag()
