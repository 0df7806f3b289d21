"""Writes tests/golden/shell_golden.npz from the reference's own thin-shell
kernels (oracle/_ref/libphysim_ref.so, built by `make -C oracle` from the
reference checkout; this script runs only where that exists).

Kernels: DiscreteShellsBending_{Energy,Gradient,HessianSPD,Shape,Dihedral}.mcg
(hinge bending, as EnergyElement_DiscreteShells.cpp:56-198 calls them) and
FEMMem_StVK_{Energy,Gradient,Hessian}.mcg (2D StVK over a row-major 3x2 F,
MaterialModels.h:376-440).

Stored per hinge case: x (4 nodes), kB, theta0, shape0 -> E, g (12), H_spd
(144), shape(x), dihedral(x).  Per membrane case: Fv (6), lame1, lame2 ->
Psi, dPsi/dFv (6), d2Psi/dFv2 (36).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "..", "oracle", "_ref", "libphysim_ref.so")


def main():
    L = ctypes.CDLL(REF)
    dp = ctypes.POINTER(ctypes.c_double)
    d = ctypes.c_double
    L.ref_ds_energy.restype = d
    L.ref_ds_energy.argtypes = [dp, dp, dp, dp, d, d, d]
    L.ref_ds_gradient.argtypes = [dp, dp, dp, dp, d, d, d, dp]
    L.ref_ds_hessian_spd.argtypes = [dp, dp, dp, dp, d, d, d, dp]
    L.ref_ds_shape.restype = d
    L.ref_ds_shape.argtypes = [dp, dp, dp, dp]
    L.ref_ds_dihedral.restype = d
    L.ref_ds_dihedral.argtypes = [dp, dp, dp, dp]
    L.ref_mem_stvk_energy.restype = d
    L.ref_mem_stvk_energy.argtypes = [dp, d, d]
    L.ref_mem_stvk_pk1.argtypes = [dp, d, d, dp]
    L.ref_mem_stvk_dpdf.argtypes = [dp, d, d, dp]
    P = lambda a: a.ctypes.data_as(dp)
    rng = np.random.default_rng(20261205)

    # hinges: a flat-ish hinge (edge along x, wings in the xz-plane) folded by
    # up to +-2.5 rad, randomly rotated, scaled and perturbed
    nh = 48
    h_x = np.zeros((nh, 12))
    h_p = np.zeros((nh, 3))
    h_E = np.zeros(nh)
    h_g = np.zeros((nh, 12))
    h_H = np.zeros((nh, 144))
    h_shape = np.zeros(nh)
    h_theta = np.zeros(nh)
    for i in range(nh):
        phi = rng.uniform(-2.5, 2.5)
        s = rng.uniform(0.002, 0.05)
        x0 = np.zeros(3)
        x1 = np.array([1.0, 0.0, 0.0])
        x2 = np.array([rng.uniform(0.2, 0.8), 0.0, rng.uniform(0.5, 1.2)])
        w = np.array([rng.uniform(0.2, 0.8), 0.0, -rng.uniform(0.5, 1.2)])
        c, sn = np.cos(phi), np.sin(phi)
        x3 = np.array([w[0], -w[2] * sn, w[2] * c])  # wing rotated about the edge
        X = np.stack([x0, x1, x2, x3]) + 0.05 * rng.standard_normal((4, 3))
        Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        X = s * X @ Q.T + rng.standard_normal(3)
        h_x[i] = X.reshape(-1)
        kB, th0, sh0 = rng.uniform(1e-6, 1e-3), rng.uniform(-0.5, 0.5), rng.uniform(0.1, 2.0)
        h_p[i] = (kB, th0, sh0)
        xs = [np.ascontiguousarray(X[k]) for k in range(4)]
        h_E[i] = L.ref_ds_energy(*[P(v) for v in xs], kB, th0, sh0)
        g = np.zeros(12)
        H = np.zeros(144)
        L.ref_ds_gradient(*[P(v) for v in xs], kB, th0, sh0, P(g))
        L.ref_ds_hessian_spd(*[P(v) for v in xs], kB, th0, sh0, P(H))
        h_g[i], h_H[i] = g, H
        h_shape[i] = L.ref_ds_shape(*[P(v) for v in xs])
        h_theta[i] = L.ref_ds_dihedral(*[P(v) for v in xs])

    # membranes: F = R (orthonormal 3x2 columns) + perturbation
    nm = 32
    m_F = np.zeros((nm, 6))
    m_l = np.zeros((nm, 2))
    m_psi = np.zeros(nm)
    m_P = np.zeros((nm, 6))
    m_C = np.zeros((nm, 36))
    for i in range(nm):
        Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        F = Q[:, :2] + 0.2 * rng.standard_normal((3, 2))
        m_F[i] = F.reshape(-1)  # row-major 3x2
        m_l[i] = rng.uniform(1e4, 1e6, 2)
        f = np.ascontiguousarray(m_F[i])
        m_psi[i] = L.ref_mem_stvk_energy(P(f), *m_l[i])
        p = np.zeros(6)
        c = np.zeros(36)
        L.ref_mem_stvk_pk1(P(f), *m_l[i], P(p))
        L.ref_mem_stvk_dpdf(P(f), *m_l[i], P(c))
        m_P[i], m_C[i] = p, c

    np.savez_compressed(os.path.join(HERE, "shell_golden.npz"), h_x=h_x, h_p=h_p, h_E=h_E,
                        h_g=h_g, h_H=h_H, h_shape=h_shape, h_theta=h_theta, m_F=m_F, m_l=m_l,
                        m_psi=m_psi, m_P=m_P, m_C=m_C)
    print("wrote", nh, "hinge and", nm, "membrane cases")


if __name__ == "__main__":
    main()
