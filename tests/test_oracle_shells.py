"""The FP64 CPU thin-shell oracle (oracle/shells.c + the mixed model of
oracle.c), pinned against the reference's own discrete-shell and 2D StVK
Maple kernels (golden vectors), closed forms of the membrane and hinge
energies, finite differences and brute-force assembly."""
import os

import numpy as np
import pytest

from oracle import orc
from physim_amd import scenes

GOLD = os.path.join(os.path.dirname(__file__), "golden", "shell_golden.npz")


def test_hinge_matches_reference_kernels():
    """E, g, SPD Hessian, shape factor and dihedral of
    DiscreteShellsBending_*.mcg, hinges folded up to +-2.5 rad."""
    g = np.load(GOLD)
    for i in range(len(g["h_E"])):
        x = g["h_x"][i]
        kB, th0, sh0 = g["h_p"][i]
        E, gr, H = orc.hinge_egh(x, kB, th0, sh0)
        assert abs(E - g["h_E"][i]) <= 1e-11 * abs(g["h_E"][i]), i
        assert np.abs(gr - g["h_g"][i]).max() <= 1e-11 * np.abs(g["h_g"][i]).max(), i
        assert np.abs(H.ravel() - g["h_H"][i]).max() <= 1e-11 * np.abs(g["h_H"][i]).max(), i
        assert abs(orc.hinge_theta(x) - g["h_theta"][i]) <= 1e-12, i
        assert abs(orc.hinge_shape(x) - g["h_shape"][i]) <= 1e-12 * g["h_shape"][i], i


def test_membrane_material_matches_reference_kernels():
    g = np.load(GOLD)
    for i in range(len(g["m_psi"])):
        F = g["m_F"][i]
        l1, l2 = g["m_l"][i]
        assert abs(orc.mem_psi(F, l1, l2) - g["m_psi"][i]) <= 1e-12 * abs(g["m_psi"][i]), i
        assert np.abs(orc.mem_pk1(F, l1, l2) - g["m_P"][i]).max() <= \
            1e-12 * np.abs(g["m_P"][i]).max(), i
        assert np.abs(orc.mem_dpdf(F, l1, l2).ravel() - g["m_C"][i]).max() <= \
            1e-12 * np.abs(g["m_C"][i]).max(), i


def _rand_tri(rng):
    X = rng.standard_normal((3, 3))
    return X


def test_membrane_rest_and_rigid_motion_are_stress_free():
    rng = np.random.default_rng(1)
    for _ in range(10):
        X = _rand_tri(rng)
        D, A = orc.membrane_rest(X)
        assert A == pytest.approx(0.5 * np.linalg.norm(np.cross(X[1] - X[0], X[2] - X[0])))
        Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        x = X @ Q.T + rng.standard_normal(3)
        E, g, H = orc.membrane_egh(x, D, A * 1e-3, 5e5, 3e5)
        assert abs(E) < 1e-12 and np.abs(g).max() < 1e-9


def test_membrane_uniaxial_stretch_closed_form():
    """In-plane stretch s along the first rest edge direction:
    E00 = ((1+s)^2 - 1)/2, others 0 -> Psi = (lambda/2 + mu) E00^2."""
    X = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.3, 0.8, 0]])
    D, A = orc.membrane_rest(X)
    s, l1, l2, t = 0.07, 4e5, 2.5e5, 2e-3
    x = X.copy()
    x[:, 0] *= 1 + s
    E, _, _ = orc.membrane_egh(x, D, A * t, l1, l2)
    e00 = ((1 + s) ** 2 - 1) / 2
    assert E == pytest.approx(A * t * (l1 / 2 + l2) * e00 ** 2, rel=1e-12)


def test_membrane_finite_differences():
    rng = np.random.default_rng(2)
    X = _rand_tri(rng)
    D, A = orc.membrane_rest(X)
    x0 = X + 0.05 * rng.standard_normal((3, 3))
    E, g, H = orc.membrane_egh(x0, D, A * 1e-3, 5e5, 3e5)
    eps = 1e-6
    gfd = np.zeros(9)
    Hfd = np.zeros((9, 9))
    xv = x0.reshape(-1)
    for k in range(9):
        xp, xm = xv.copy(), xv.copy()
        xp[k] += eps
        xm[k] -= eps
        Ep, gp, _ = orc.membrane_egh(xp, D, A * 1e-3, 5e5, 3e5)
        Em, gm, _ = orc.membrane_egh(xm, D, A * 1e-3, 5e5, 3e5)
        gfd[k] = (Ep - Em) / (2 * eps)
        Hfd[:, k] = (gp - gm) / (2 * eps)
    assert np.linalg.norm(g - gfd) <= 1e-6 * np.linalg.norm(g)
    assert np.linalg.norm(H - Hfd) <= 1e-6 * np.linalg.norm(H)


def test_hinge_fold_closed_form():
    """A flat hinge folded by phi about its edge: theta = +-phi, and the
    stored energy is kB shape0 phi^2 with the flat rest shape."""
    X = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.4, 0, 0.9], [0.6, 0, -0.7]])
    th0 = orc.hinge_theta(X)
    sh0 = orc.hinge_shape(X)
    assert abs(th0) < 1e-15
    assert sh0 == pytest.approx(1.0 / (4 * 0.9 + 4 * 0.7))
    for phi in (0.1, -0.7, 1.9):
        x = X.copy()
        c, s = np.cos(phi), np.sin(phi)
        x[3] = [0.6, 0.7 * s, -0.7 * c]
        assert abs(abs(orc.hinge_theta(x)) - abs(phi)) < 1e-12
        E, _, _ = orc.hinge_egh(x, 2.0, th0, sh0)
        assert E == pytest.approx(2.0 * sh0 * phi ** 2, rel=1e-12)


def test_hinge_gradient_finite_differences():
    rng = np.random.default_rng(3)
    for _ in range(5):
        x = rng.standard_normal(12)
        d = orc.hinge_dtheta(x)
        eps = 1e-6
        fd = np.zeros(12)
        for k in range(12):
            xp, xm = x.copy(), x.copy()
            xp[k] += eps
            xm[k] -= eps
            fd[k] = (orc.hinge_theta(xp) - orc.hinge_theta(xm)) / (2 * eps)
        assert np.abs(d - fd).max() <= 1e-7 * np.abs(d).max()


def test_cloth_mesh_counts():
    """create2DGridIndex_Tri3 / interior edges: 2 (N-1)^2 triangles and
    (N-1)(3N-5) hinges; every hinge's nodes span two distinct triangles."""
    for n in (3, 6, 11):
        s = scenes.cloth(n)
        assert len(s.F) == 2 * (n - 1) ** 2
        assert len(s.hinges) == (n - 1) * (3 * n - 5)
        assert np.all(s.hinges[:, 2] != s.hinges[:, 3])


def _cloth_oracle(s):
    T, kind, l1, l2, rho, extra = s.oracle_arrays()
    M = orc.Model(s.X0, T, kind, l1, l2, rho, extra)
    s.apply(M)
    return M


def test_cloth_mass_lumps_density_times_area():
    s = scenes.cloth(5, size=0.8, density=0.3)
    M = _cloth_oracle(s)
    mass = M.mass().reshape(-1, 3)[:, 0]
    assert mass.sum() == pytest.approx(0.3 * 0.8 * 0.8, rel=1e-12)


def test_cloth_assembly_matches_dense_sum():
    """The lower CSR Hessian equals the dense sum of element matrices
    (triangle 9x9 + hinge 12x12) with the fixed-DoF identity projection."""
    s = scenes.cloth(4, jitter=0.1, seed=1)
    M = _cloth_oracle(s)
    rng = np.random.default_rng(4)
    x = s.X0 + 0.01 * rng.standard_normal(s.X0.shape)
    M.set_x(x)
    rp, ci, v = M.hessian_lower()
    A = orc.lower_to_dense(rp, ci, v)
    N = 3 * s.n
    B = np.zeros((N, N))
    nt = len(s.F)
    for t, f in enumerate(s.F):
        D, Ar = orc.membrane_rest(s.X0[f])
        _, _, H = orc.membrane_egh(x[f], D, Ar * s.thickness, s.lame1, s.lame2)
        dof = (3 * f[:, None] + np.arange(3)).ravel()
        B[np.ix_(dof, dof)] += H
    for hh in s.hinges:
        th0 = orc.hinge_theta(s.X0[hh])
        sh0 = orc.hinge_shape(s.X0[hh])
        _, _, H = orc.hinge_egh(x[hh], s.bending_k, th0, sh0)
        dof = (3 * hh[:, None] + np.arange(3)).ravel()
        B[np.ix_(dof, dof)] += H
    fx = s.fixed.reshape(-1).astype(bool)
    B[fx, :] = 0
    B[:, fx] = 0
    B[fx, fx] = 1
    assert np.abs(A - B).max() <= 1e-12 * np.abs(B).max()
    assert nt > 0


def test_cloth_newton_converges():
    """A 9x9 sheet pinned at two corners sags under gravity to |g| <= 1e-8
    with the reference's Newton (one load stage per iteration)."""
    s = scenes.cloth(9, size=0.2, stages=10)
    M = _cloth_oracle(s)
    r = M.newton(tolMaxError=1e-8, tolMinImpr=0.0, tolMinStep=0.0, numMaxIters=200,
                 ls_noise_guard=1e-12)
    assert r["result"] == "SUCCESS", r
    x = M.get_x().reshape(-1, 3)
    assert x[:, 1].min() < -1e-3


def test_batch_beam_traction_lumping_closed_form():
    """The lumped traction of a batch beam sums to traction x face area."""
    rng = np.random.default_rng(9)
    for _ in range(5):
        s = scenes.batch_beam(rng, cell=0.01, dmin=(2, 2, 4), dmax=(6, 6, 12))
        f = s.forces.sum(0)
        loaded = np.abs(s.forces).sum(1) > 0
        if not loaded.any():
            continue
        X = s.X0[loaded]
        ext = X.max(0) - X.min(0)
        area = np.prod(ext[ext > 1e-12])
        t = np.linalg.norm(f) / area
        assert 0.0 <= t <= 50.0
        # every loaded node carries the same direction
        d = s.forces[loaded] / np.linalg.norm(s.forces[loaded], axis=1)[:, None]
        assert np.abs(d - d[0]).max() < 1e-12
        assert s.fixed.any()
