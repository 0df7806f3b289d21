"""The oracle's implicit-dynamics problem (OptimProblem_BasicDynamic.cpp:
20-170, one backward-Euler step per SolveFull), pinned by what the
mathematics of the step fixes: free fall under constant gravity is exact,
internal forces conserve linear momentum, backward Euler dissipates the total
energy of an elastic body, a very long step reduces to the static
equilibrium, and the step's gradient and Hessian are the derivatives of its
objective."""
import numpy as np
import pytest

from oracle import orc
from tests.helpers import oracle_model, small_beam

# tolMinStep / tolMinImpr 0 and the noise guard: the last Newton steps of these tiny
# problems change the energy by less than its FP64 rounding
SOLVE = dict(tolMaxError=1e-9, lsTolMaxError=1e-12, lsNumMaxIters=5000, tolMinStep=0.0,
             tolMinImpr=0.0,
             ls_noise_guard=1e-12)


def _free(scene):
    """The scene without any clamped node."""
    scene.fixed = np.zeros_like(scene.fixed)
    return scene


def test_free_fall_is_exact():
    """Backward Euler under a constant load with no strain: v_n = n dt g and
    x_n = X0 + dt^2 g n (n + 1) / 2 for every node."""
    s = _free(small_beam(dims=(2, 2, 4)))
    s.stages = 1
    M = oracle_model(s)
    D = orc.Dynamic(M)
    dt = 0.01
    g = np.array(s.gravity)
    for n in range(1, 4):
        r = D.step(dt, **SOLVE)
        assert r["result"] == "SUCCESS", r
        x = M.get_x().reshape(-1, 3)
        v = D.get_velocity().reshape(-1, 3)
        want_x = s.X0 + dt * dt * g * n * (n + 1) / 2
        scale = dt * dt * np.abs(g).max() * n * (n + 1) / 2
        assert np.abs(x - want_x).max() <= 1e-8 * scale, n
        assert np.abs(v - n * dt * g).max() <= 1e-8 * n * dt * np.abs(g).max(), n


def test_linear_momentum_is_conserved():
    """No load and no clamp: sum_a g_a = 0 per element, so M v is constant."""
    s = _free(small_beam(dims=(2, 2, 5)))
    s.gravity = (0.0, 0.0, 0.0)
    M = oracle_model(s)
    D = orc.Dynamic(M)
    rng = np.random.default_rng(2)
    c = s.X0.mean(0)
    v0 = np.cross([0.0, 0.0, 3.0], s.X0 - c) + [0.1, -0.2, 0.05]
    v0 += 0.05 * rng.standard_normal(v0.shape)
    D.set_velocity(v0)
    m = M.mass().reshape(-1, 3)
    P0 = (m * v0).sum(0)
    for _ in range(5):
        r = D.step(0.005, **SOLVE)
        assert r["result"] == "SUCCESS", r
        P = (m * D.get_velocity().reshape(-1, 3)).sum(0)
        assert np.abs(P - P0).max() <= 1e-9 * np.abs(P0).max()


def test_backward_euler_dissipates_energy():
    """Clamped beam released with a transverse velocity, no gravity: the total
    energy 1/2 v.M v + E_elastic never grows from one step to the next."""
    s = small_beam(dims=(2, 2, 6))
    s.gravity = (0.0, 0.0, 0.0)
    M = oracle_model(s)
    D = orc.Dynamic(M)
    v0 = np.zeros_like(s.X0)
    v0[:, 1] = 2.0 * s.X0[:, 2] / s.X0[:, 2].max()
    v0[s.fixed.astype(bool)] = 0.0
    D.set_velocity(v0)
    m = M.mass()
    Hs = [0.5 * v0.reshape(-1) @ (m * v0.reshape(-1)) + M.energy()]
    for _ in range(8):
        r = D.step(0.002, **SOLVE)
        assert r["result"] == "SUCCESS", r
        v = D.get_velocity()
        Hs.append(0.5 * v @ (m * v) + M.energy())
    Hs = np.array(Hs)
    assert np.all(np.diff(Hs) <= 1e-9 * Hs[0]), Hs
    assert Hs[-1] < 0.99 * Hs[0]


def test_long_step_is_the_static_equilibrium():
    """dt -> infinity removes the inertia term: one step from rest under the
    full load lands on the static solution (stiff beam, small deflection)."""
    s = small_beam(dims=(2, 2, 5))
    s.lame1 = s.lame1 * 100.0
    s.lame2 = s.lame2 * 100.0
    s.stages = 1
    Ms = oracle_model(s)
    rs = Ms.newton(**SOLVE)
    assert rs["result"] == "SUCCESS", rs
    Md = oracle_model(s)
    D = orc.Dynamic(Md)
    rd = D.step(1e6, **SOLVE)
    assert rd["result"] == "SUCCESS", rd
    diff = np.abs(Md.get_x() - Ms.get_x()).max()
    assert diff <= 1e-9 * s.bbox_diag()
    assert np.abs(Ms.get_x() - s.X0.reshape(-1)).max() > 1e-5  # it did deflect


def test_step_objective_derivatives():
    """Central differences of the step's objective (eps 1e-6, the reference's
    element check) give its gradient on the free DoFs; differences of the
    gradient give its Hessian."""
    s = small_beam(dims=(2, 2, 3))
    M = oracle_model(s)
    D = orc.Dynamic(M)
    rng = np.random.default_rng(5)
    D.set_velocity(0.3 * rng.standard_normal(s.X0.shape))
    D.begin(0.01)
    free = ~s.fixed.astype(bool).reshape(-1)
    x = M.get_x() + 0.002 * rng.standard_normal(M.N) * free
    M.set_x(x)
    g = D.gradient()
    assert np.all(g[~free] == 0.0)
    eps = 1e-6
    fd = np.zeros(M.N)
    for k in np.flatnonzero(free):
        xp = x.copy(); xp[k] += eps
        M.set_x(xp); ep = D.energy()
        xm = x.copy(); xm[k] -= eps
        M.set_x(xm); em = D.energy()
        fd[k] = (ep - em) / (2 * eps)
    M.set_x(x)
    assert np.abs(fd[free] - g[free]).max() <= 1e-6 * np.abs(g[free]).max()
    rp, ci, v = D.hessian_lower()
    H = orc.lower_to_dense(rp, ci, v)
    d = rng.standard_normal(M.N) * free
    M.set_x(x + eps * d); gp = D.gradient()
    M.set_x(x - eps * d); gm = D.gradient()
    M.set_x(x)
    Hd = (gp - gm) / (2 * eps)
    assert np.abs((H @ d)[free] - Hd[free]).max() <= 1e-6 * np.abs(Hd).max()
    fixed = np.flatnonzero(~free)
    assert np.all(H[np.ix_(fixed, fixed)] == np.eye(fixed.size))


def test_inertia_pieces_closed_form():
    """At the predictor x = x0 + dt v0 the inertia term and its gradient
    vanish and the Hessian gains exactly M / dt^2 on its diagonal."""
    s = small_beam(dims=(2, 2, 3))
    M = oracle_model(s)
    D = orc.Dynamic(M)
    rng = np.random.default_rng(6)
    v0 = 0.2 * rng.standard_normal(s.X0.shape)
    v0[s.fixed.astype(bool)] = 0.0
    D.set_velocity(v0)
    dt = 0.02
    D.begin(dt)
    x = M.get_x()
    np.testing.assert_allclose(x, s.X0.reshape(-1) + dt * v0.reshape(-1), rtol=0, atol=1e-15)
    assert D.energy() == pytest.approx(M.energy(), rel=1e-14, abs=1e-300)
    np.testing.assert_allclose(D.gradient(), M.gradient(), rtol=0, atol=1e-12)
    rp, ci, hv = D.hessian_lower()
    _, _, sv = M.hessian_lower()
    m = M.mass()
    free = ~s.fixed.astype(bool).reshape(-1)
    diag = np.array([p for r in range(M.N) for p in range(rp[r], rp[r + 1]) if ci[p] == r])
    extra = np.zeros_like(hv)
    extra[diag] = (m / dt ** 2) * free
    np.testing.assert_allclose(hv, sv + extra, rtol=1e-14, atol=1e-9)
