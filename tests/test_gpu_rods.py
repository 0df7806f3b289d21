"""libpgn discrete-elastic-rod path on the GPU vs the FP64 rod oracle, through
the C ABI.

Tolerances: element and assembled quantities are FP64 on both sides:
relative 1e-10 (normwise); frames and reference twists absolute 1e-12;
converged positions within 1e-6 x each rod's bounding-box diagonal.
"""
import os
import sys

import numpy as np
import pytest

import physim_amd as pg
from oracle import orc
from physim_amd import scenes

pytestmark = pytest.mark.gpu

ROD_OPTS = dict(tolMaxError=1e-8, numMaxIters=100, tolMinImpr=0.0, tolMinStep=0.0)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def _pair(ctx, s, gravity=None, precision=pg.FP64):
    G = pg.Rods(ctx, s.rods, s.frame0_normals(), s.young, s.shear, s.radius, s.radius,
                s.density, precision=precision)
    mask = s.fixed_mask()
    G.set_fixed(mask)
    O = []
    off = G.offsets
    for r, X in enumerate(s.rods):
        R = orc.Rod(X, s.frame0_normals()[r], s.young, s.shear, s.radius, s.radius, s.density)
        R.set_fixed(mask[off[r]:off[r + 1]])
        O.append(R)
    if gravity is not None:
        G.set_gravity(gravity, 1, g_ini=gravity)
        for R in O:
            R.set_gravity(gravity, 1, g_ini=gravity)
    return G, O


def _perturb(s, G, amp, seed):
    rng = np.random.default_rng(seed)
    x = G.get_x()
    mask = s.fixed_mask().astype(bool)
    x = x + amp * rng.standard_normal(x.size) * ~mask
    return x


def test_rods_energy_gradient_hessian_match_oracle(ctx):
    _say("rods eval vs oracle")
    s = scenes.rod_batch(5, n_min=6, n_max=40, seed=11, density=3e-5)
    G, O = _pair(ctx, s, gravity=[0.0, -9.81, 0.0])
    x = _perturb(s, G, 2e-3, 1)
    G.set_x(x)
    off = G.offsets
    for r, R in enumerate(O):
        R.set_x(x[off[r]:off[r + 1]])
    E = G.energy()
    g = G.gradient()
    for r, R in enumerate(O):
        assert abs(E[r] - R.energy()) <= 1e-10 * abs(R.energy()), r
        assert _rel(g[off[r]:off[r + 1]], R.gradient()) <= 1e-10, r
        H = G.hessian_dense(r)
        Ho = R.hessian()
        assert _rel(H, Ho) <= 1e-10, r
        assert np.array_equal(H, H.T), r


def test_rods_element_bt_matches_oracle(ctx):
    _say("rods bend-twist elements vs oracle")
    s = scenes.rod_batch(3, n_min=5, n_max=25, seed=12, density=3e-5)
    G, O = _pair(ctx, s)
    x = _perturb(s, G, 3e-3, 2)
    G.set_x(x)
    E, g8, H = G.element_bt()
    off = G.offsets
    v0 = 0
    for r, R in enumerate(O):
        R.set_x(x[off[r]:off[r + 1]])
        n = R.n
        Eo_sum = 0.0
        assert np.all(g8[v0] == 0.0) and np.all(H[v0 + n - 1] == 0.0)
        assert E[v0] == 0.0 and E[v0 + n - 1] == 0.0  # no element at the ends
        # element energies sum to the rod's bend-twist energy: total minus stretch
        X = x[off[r]:off[r] + 3 * n].reshape(-1, 3)
        for e in range(n - 1):
            L0 = np.linalg.norm(s.rods[r][e + 1] - s.rods[r][e])
            Eo_sum += orc.rod_stretch(X[e + 1] - X[e], L0, s.young * np.pi * s.radius ** 2)[0]
        bt = R.energy() - Eo_sum
        assert abs(E[v0:v0 + n].sum() - bt) <= 1e-10 * abs(bt), r
        v0 += n


def test_rods_update_kinematics_matches_oracle(ctx):
    _say("rods kinematics vs oracle")
    s = scenes.rod_batch(4, n_min=5, n_max=30, seed=13, density=3e-5)
    G, O = _pair(ctx, s)
    off = G.offsets
    for step in range(3):
        x = _perturb(s, G, 2e-3, 10 + step)
        G.set_x(x)
        G.update_kinematics()
        for r, R in enumerate(O):
            R.set_x(x[off[r]:off[r + 1]])
            R.update_kinematics()
    t, nn, rt = G.frames()
    e0 = v0 = 0
    for r, R in enumerate(O):
        to, no, rto = R.frames()
        ne = R.n - 1
        assert np.abs(t[e0:e0 + ne] - to).max() <= 1e-12, r
        assert np.abs(nn[e0:e0 + ne] - no).max() <= 1e-12, r
        assert np.abs(rt[v0:v0 + R.n] - rto).max() <= 1e-12, r
        e0 += ne
        v0 += R.n


GOLD = os.path.join(os.path.dirname(__file__), "golden", "rod_golden.npz")


@pytest.mark.parametrize("precision,tol", [(pg.FP64, 1e-10), (pg.MIXED, 1e-5)])
def test_rod_bt_kernel_matches_reference_golden(ctx, precision, tol):
    """The device bend-twist kernel (FP64, and FP32 as the mixed batches use
    it for their Hessians) on the inputs of the golden vectors written from
    the reference's own getBT{Energy,Gradient,Hessian}OPT kernels: E, g and
    H normwise within 1e-10 (FP64) / 1e-5 (FP32) relative."""
    _say(f"rod bend-twist kernel vs golden precision={precision}")
    g = np.load(GOLD)
    E, gr, H = pg.rod_bt_eval(ctx, g["b_prm"], g["b_e1"], g["b_e2"], g["b_th"], precision)
    for i in range(len(E)):
        assert abs(E[i] - g["b_E"][i]) <= tol * max(abs(g["b_E"][i]), 1e-300), i
        assert _rel(gr[i], g["b_g"][i]) <= tol, i
        assert _rel(H[i], g["b_H"][i].reshape(8, 8)) <= tol, i
        assert np.array_equal(H[i], H[i].T), i


def test_rods_mixed_element_hessians_and_newton(ctx):
    """A PGN_MIXED rod batch: FP32 bend-twist Hessians within 1e-5 of the
    FP64 oracle's, FP64 energies and gradients within 1e-10, and Newton
    converged to the oracle's positions within 1e-6 x bbox diagonal."""
    _say("rods mixed precision vs oracle")
    s = scenes.rod_batch(6, n_min=10, n_max=60, seed=14, density=3e-6)
    G, O = _pair(ctx, s, precision=pg.MIXED)
    G64, _ = _pair(ctx, s)
    x = _perturb(s, G, 2e-3, 3)
    G.set_x(x)
    G64.set_x(x)
    E, g8, H = G.element_bt()
    E6, g6, H6 = G64.element_bt()
    assert _rel(E, E6) <= 1e-12 and _rel(g8, g6) <= 1e-12
    assert _rel(H, H6) <= 1e-5
    G.reset()
    G.set_gravity([0.0, -9.81, 0.0], 10)
    for R in O:
        R.set_gravity([0.0, -9.81, 0.0], 10)
    res = G.newton(**ROD_OPTS)
    x = G.get_x()
    off = G.offsets
    diag = s.bbox_diags()
    for r, R in enumerate(O):
        ro = R.newton(**ROD_OPTS)
        assert ro["result"] == "SUCCESS", (r, ro)
        assert res[r]["result"] == "SUCCESS", (r, res[r])
        n = R.n
        err = np.abs(x[off[r]:off[r] + 3 * n] - R.get_x()[:3 * n]).max()
        assert err <= 1e-6 * diag[r], (r, err)


def test_rods_newton_matches_oracle(ctx):
    """Every rod of the batch keeps its own Newton state (status machine,
    line search, load stage); the converged positions match the oracle's."""
    _say("rods newton vs oracle")
    s = scenes.rod_batch(6, n_min=10, n_max=60, seed=14, density=3e-6)
    G, O = _pair(ctx, s)
    G.set_gravity([0.0, -9.81, 0.0], 10)
    for R in O:
        R.set_gravity([0.0, -9.81, 0.0], 10)
    res = G.newton(**ROD_OPTS)
    x = G.get_x()
    off = G.offsets
    diag = s.bbox_diags()
    for r, R in enumerate(O):
        ro = R.newton(**ROD_OPTS)
        assert ro["result"] == "SUCCESS", (r, ro)
        assert res[r]["result"] == "SUCCESS", (r, res[r])
        n = R.n
        err = np.abs(x[off[r]:off[r] + 3 * n] - R.get_x()[:3 * n]).max()
        assert err <= 1e-6 * diag[r], (r, err)


def test_rods_batch_result_independent_of_batch(ctx):
    """A rod solved alone and inside a batch reaches the same positions."""
    _say("rods alone vs batch")
    s = scenes.rod_batch(4, n_min=10, n_max=50, seed=15, density=3e-6)
    G, _ = _pair(ctx, s)
    G.set_gravity([0.0, -9.81, 0.0], 10)
    G.newton(**ROD_OPTS)
    xb = G.get_x()
    one = scenes.RodScene([s.rods[2]], s.young, s.shear, s.radius, s.density, s.stages, "one")
    G1, _ = _pair(ctx, one)
    G1.set_gravity([0.0, -9.81, 0.0], 10)
    G1.newton(**ROD_OPTS)
    x1 = G1.get_x()
    off = G.offsets
    n = len(s.rods[2])
    assert np.abs(xb[off[2]:off[3]][:3 * n] - x1[:3 * n]).max() <= 1e-6 * s.bbox_diags()[2]
