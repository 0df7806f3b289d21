"""The numpy restatement of the two-level preconditioner (oracle/twolevel.py),
pinned by what fixes it independently of the device code:

- an unstressed, unconstrained elastic body has the six rigid-body motions
  as null vectors of its Hessian, so A P = 0 for one aggregate spanning the
  mesh (a wrong component or offset in P breaks it);
- the binning rule's box counts and node boxes in closed form on a grid beam;
- dependent columns are dropped (a fixed aggregate keeps none, a line of free
  nodes loses the rotation about its own axis);
- the preconditioner is symmetric positive definite and cuts the CG's
  iterations on a loaded beam's tangent.
"""
import numpy as np

from oracle import orc
from oracle import twolevel as tl
from physim_amd import scenes
from tests.helpers import oracle_model, small_beam


def _dense_hessian(O):
    return orc.lower_to_dense(*O.hessian_lower())


def test_rigid_modes_annihilate_unstressed_hessian():
    s = small_beam(dims=(2, 2, 4))
    O = orc.Model(s.X0, s.T, s.material, s.lame1, s.lame2, s.density)  # nothing fixed
    O.set_x(s.X0)
    A = _dense_hessian(O)
    agg = np.zeros(s.n, dtype=np.int32)  # one aggregate spanning the mesh
    fixed = np.zeros_like(s.fixed)
    cm = tl.kept_columns(s.X0, agg, fixed)
    assert cm[0] == 0x3F
    P, keep = tl.prolongation(s.X0, agg, fixed, cm)
    assert keep.all()
    assert np.abs(A @ P).max() <= 1e-10 * np.abs(A).max()
    # a deformed (stressed) state: the rotations are no longer null vectors
    rng = np.random.default_rng(1)
    O.set_x(s.X0 + 0.02 * rng.standard_normal(s.X0.shape))
    A2 = _dense_hessian(O)
    assert np.abs(A2 @ P[:, :3]).max() <= 1e-10 * np.abs(A2).max()  # translations still
    assert np.abs(A2 @ P[:, 3:]).max() > 1e-6 * np.abs(A2).max()


def test_bins_closed_form_on_grid_beam():
    s = scenes.config1()  # 0.1 x 0.1 x 0.4 m, 10 x 10 x 40 cells
    b = tl.bins_for(s.X0, 128)
    # cbrt(128 / 0.004) = 31.75 boxes per metre -> 3.17, 3.17, 12.7
    assert b == [3, 3, 13]
    agg = tl.aggregate(s.X0, 128)
    assert agg.max() + 1 == 3 * 3 * 13
    ijk = np.rint(s.X0 / np.array([0.01, 0.01, 0.01])).astype(int)  # grid indices
    box = np.minimum(ijk * np.array(b) // np.array([10, 10, 40]), np.array(b) - 1)
    want = (box[:, 2] * 3 + box[:, 1]) * 3 + box[:, 0]
    assert np.array_equal(agg, want)
    # a flat sheet: one box across its thickness
    X = np.stack(np.meshgrid(np.linspace(0, 1, 9), np.linspace(0, 1, 9), [0.0]), -1).reshape(-1, 3)
    assert tl.bins_for(X, 64)[2] == 1


def test_dependent_columns_dropped():
    # aggregate 0: four free nodes on a line along x; aggregate 1: all fixed;
    # aggregate 2: a free tetrahedron's corners
    X = np.array([[0, 0, 0], [0.1, 0, 0], [0.2, 0, 0], [0.3, 0, 0],
                  [5, 0, 0], [5.1, 0, 0],
                  [10, 0, 0], [10.1, 0, 0], [10, 0.1, 0], [10, 0, 0.1]], dtype=float)
    agg = np.array([0, 0, 0, 0, 1, 1, 2, 2, 2, 2], dtype=np.int32)
    fixed = np.zeros((10, 3), dtype=np.uint8)
    fixed[4:6] = 1
    cm = tl.kept_columns(X, agg, fixed)
    assert cm[0] == 0x3F & ~(1 << 3)  # no rotation about the line's own axis
    assert cm[1] == 0
    assert cm[2] == 0x3F


def test_preconditioner_spd_and_effective():
    s = small_beam(dims=(2, 2, 8))
    O = oracle_model(s)
    r = O.newton(tolMaxError=1e-8, tolMinImpr=0.0, tolMinStep=0.0, ls_noise_guard=1e-12)
    assert r["result"] == "SUCCESS"
    A = _dense_hessian(O)
    agg = tl.aggregate(s.X0, 8)
    cm = tl.kept_columns(s.X0, agg, s.fixed)
    P, keep = tl.prolongation(s.X0, agg, s.fixed, cm)
    Ac = tl.coarse_matrix(A, P, keep)
    N = A.shape[0]
    Minv = np.stack([tl.apply(A, e, P, Ac) for e in np.eye(N)], 1)
    assert np.abs(Minv - Minv.T).max() <= 1e-10 * np.abs(Minv).max()
    assert np.linalg.eigvalsh(0.5 * (Minv + Minv.T)).min() > 0

    def pcg_iters(M, b, tol=1e-8):
        x = np.zeros_like(b)
        rr = b.copy()
        z = M(rr)
        p = z.copy()
        rz = rr @ z
        for k in range(1, 5000):
            q = A @ p
            a = rz / (p @ q)
            x += a * p
            rr -= a * q
            if np.linalg.norm(rr) <= tol * np.linalg.norm(b):
                return k
            z = M(rr)
            rz2 = rr @ z
            p = z + rz2 / rz * p
            rz = rz2
        return 5000

    b = np.random.default_rng(2).standard_normal(N)
    b[s.fixed.reshape(-1).astype(bool)] = 0.0
    k_bj = pcg_iters(lambda v: tl.block_jacobi(A, v), b)
    k_2l = pcg_iters(lambda v: tl.apply(A, v, P, Ac), b)
    assert k_2l < k_bj
