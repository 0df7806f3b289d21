"""Writes tests/golden/rod_golden.npz from the reference's own discrete-elastic-
rod kernels (oracle/_ref/libphysim_ref.so, built by `make -C oracle` from the
reference checkout; this script runs only where that exists).

Kernels: RodEdge3D{Energy,Force,Jacobian}OPT.mcg (stretch, as called by
EnergyElement_DER_Stretch.cpp:53-120), getBT{Energy,Gradient,Hessian}OPT.h
(bend-twist, as called by EnergyElement_DER_BendTwist.cpp:67-249) and
computeReferenceTwistChange.h (Mesh_DER.cpp:337-353).

Stored per case: the inputs in the kernels' own argument order and the
outputs they write.  Bend-twist inputs are sampled from rod-like states:
reference frames close to the current tangents (as they are after a
Newton step), moderate bend and twist.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "..", "oracle", "_ref", "libphysim_ref.so")


def _frame(rng, t):
    t = t / np.linalg.norm(t)
    v = rng.standard_normal(3)
    n = v - (v @ t) * t
    return t, n / np.linalg.norm(n)


def main():
    L = ctypes.CDLL(REF)
    dp = ctypes.POINTER(ctypes.c_double)
    d = ctypes.c_double
    L.ref_rod_stretch_energy.restype = d
    L.ref_rod_stretch_energy.argtypes = [dp, d, d, d, d]
    L.ref_rod_stretch_force.argtypes = [dp, d, d, d, d, dp]
    L.ref_rod_stretch_jacobian.argtypes = [dp, d, d, d, d, dp]
    L.ref_bt_energy.restype = d
    L.ref_bt_energy.argtypes = [dp, dp, dp, d, d]
    L.ref_bt_gradient.argtypes = [dp, dp, dp, d, d, dp]
    L.ref_bt_hessian.argtypes = [dp, dp, dp, d, d, dp]
    L.ref_reference_twist_change.restype = d
    L.ref_reference_twist_change.argtypes = [dp, dp, dp, dp, d]
    P = lambda a: a.ctypes.data_as(dp)
    rng = np.random.default_rng(20261204)

    # stretch: edge vectors around the rest length, compressed and stretched
    ns = 32
    s_x12 = rng.standard_normal((ns, 3)) * 0.01
    s_prm = np.stack([rng.uniform(0.005, 0.02, ns), rng.uniform(5e-4, 2e-3, ns),
                      rng.uniform(5e-4, 2e-3, ns), rng.uniform(1e6, 1e8, ns)], axis=1)
    s_E = np.zeros(ns)
    s_f = np.zeros((ns, 3))
    s_J = np.zeros((ns, 9))
    for i in range(ns):
        x = np.ascontiguousarray(s_x12[i])
        L0, rh, rw, E = s_prm[i]
        s_E[i] = L.ref_rod_stretch_energy(P(x), L0, rh, rw, E)
        f = np.zeros(3)
        J = np.zeros(9)
        L.ref_rod_stretch_force(P(x), L0, rh, rw, E, P(f))
        L.ref_rod_stretch_jacobian(P(x), L0, rh, rw, E, P(J))
        s_f[i], s_J[i] = f, J

    # bend-twist: prm = shear[2], young[2], radw[2], radh[2], L0, L1, K0[2],
    # twist0, nRef1, tRef1, nRef2, tRef2, rt (27 values)
    nb = 48
    b_prm = np.zeros((nb, 27))
    b_e1 = np.zeros((nb, 3))
    b_e2 = np.zeros((nb, 3))
    b_th = np.zeros((nb, 2))
    b_E = np.zeros(nb)
    b_g = np.zeros((nb, 8))
    b_H = np.zeros((nb, 64))
    for i in range(nb):
        e1 = rng.standard_normal(3) * 0.01
        e2 = e1 + 0.4 * np.linalg.norm(e1) * rng.standard_normal(3)
        u1, n1 = _frame(rng, e1 + 0.1 * np.linalg.norm(e1) * rng.standard_normal(3))
        u2, n2 = _frame(rng, e2 + 0.1 * np.linalg.norm(e2) * rng.standard_normal(3))
        p = b_prm[i]
        p[0:2] = rng.uniform(1e6, 1e7, 2)
        p[2:4] = rng.uniform(1e7, 3e7, 2)
        p[4:8] = rng.uniform(5e-4, 2e-3, 4)
        p[8:10] = rng.uniform(0.005, 0.02, 2)
        p[10:12] = rng.standard_normal(2) * 0.3
        p[12] = 0.2 * rng.standard_normal()
        p[13:16], p[16:19], p[19:22], p[22:25] = n1, u1, n2, u2
        p[25] = 0.3 * rng.standard_normal()
        th = rng.standard_normal(2) * 0.5
        b_e1[i], b_e2[i], b_th[i] = e1, e2, th
        pp = np.ascontiguousarray(p)
        b_E[i] = L.ref_bt_energy(P(pp), P(e1), P(e2), th[0], th[1])
        g = np.zeros(8)
        H = np.zeros(64)
        L.ref_bt_gradient(P(pp), P(e1), P(e2), th[0], th[1], P(g))
        L.ref_bt_hessian(P(pp), P(e1), P(e2), th[0], th[1], P(H))
        b_g[i], b_H[i] = g, H

    # reference twist change: frames of two adjacent edges
    nt = 32
    t_in = np.zeros((nt, 12))
    t_rt = 0.5 * rng.standard_normal(nt)
    t_out = np.zeros(nt)
    for i in range(nt):
        a = rng.standard_normal(3)
        b = a + 0.5 * rng.standard_normal(3)
        t1, n1 = _frame(rng, a)
        t2, n2 = _frame(rng, b)
        t_in[i] = np.concatenate([t1, n1, t2, n2])
        v = [np.ascontiguousarray(t_in[i, 3 * k:3 * k + 3]) for k in range(4)]
        t_out[i] = L.ref_reference_twist_change(P(v[0]), P(v[1]), P(v[2]), P(v[3]), t_rt[i])

    np.savez_compressed(os.path.join(HERE, "rod_golden.npz"),
                        s_x12=s_x12, s_prm=s_prm, s_E=s_E, s_f=s_f, s_J=s_J,
                        b_prm=b_prm, b_e1=b_e1, b_e2=b_e2, b_th=b_th, b_E=b_E, b_g=b_g,
                        b_H=b_H, t_in=t_in, t_rt=t_rt, t_out=t_out)
    print("wrote", ns, "stretch,", nb, "bend-twist,", nt, "twist cases")


if __name__ == "__main__":
    main()
