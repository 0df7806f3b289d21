"""Writes tests/golden/material_golden.npz from the reference's own Maple
material kernels (oracle/_ref/libphysim_ref.so, built by `make -C oracle`
from the reference checkout; this script runs only where that exists).

Cases: well-conditioned F, near-singular F (J -> 0+) and inverted F (J < 0),
for CoNH (FEMVol_CoNH_*.mcg) and StVK (FEMVol_StVK_*.mcg).  Stored: column-
major vec(F), lame1, lame2, Psi, P = dPsi/dvec(F), C = d2Psi/dvec(F)^2.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "..", "oracle", "_ref", "libphysim_ref.so")


def main():
    L = ctypes.CDLL(REF)
    dp = ctypes.POINTER(ctypes.c_double)
    for k in ("conh", "stvk"):
        getattr(L, f"ref_{k}_energy").restype = ctypes.c_double
        getattr(L, f"ref_{k}_energy").argtypes = [dp, ctypes.c_double, ctypes.c_double]
        for f in ("pk1", "dpdf"):
            getattr(L, f"ref_{k}_{f}").argtypes = [dp, ctypes.c_double, ctypes.c_double, dp]
    rng = np.random.default_rng(20261203)
    Fs, kinds = [], []
    for kind in (0, 1):
        cases = []
        while len(cases) < 24:  # well conditioned
            F = np.eye(3) + 0.25 * rng.standard_normal((3, 3))
            if np.linalg.det(F) > 0.2:
                cases.append(F)
        while len(cases) < 36:  # J -> 0+
            U, _, Vt = np.linalg.svd(rng.standard_normal((3, 3)))
            s = np.array([1.0 + 0.2 * rng.random(), 0.8 + 0.2 * rng.random(),
                          10.0 ** rng.uniform(-6, -2)])
            F = U @ np.diag(s) @ Vt
            if np.linalg.det(F) < 0:
                F[:, 0] *= -1
            cases.append(F)
        while len(cases) < 48:  # inverted
            F = np.eye(3) + 0.25 * rng.standard_normal((3, 3))
            F[:, 2] *= -1
            if np.linalg.det(F) < -0.05:
                cases.append(F)
        Fs += cases
        kinds += [kind] * len(cases)
    n = len(Fs)
    Fv = np.stack([F.T.reshape(-1) for F in Fs])  # column-major vec
    l1 = rng.uniform(1e3, 1e6, n)
    l2 = rng.uniform(1e3, 1e6, n)
    psi = np.zeros(n)
    P = np.zeros((n, 9))
    C = np.zeros((n, 81))
    for i in range(n):
        k = "conh" if kinds[i] == 0 else "stvk"
        f = np.ascontiguousarray(Fv[i])
        fp = f.ctypes.data_as(dp)
        psi[i] = getattr(L, f"ref_{k}_energy")(fp, l1[i], l2[i])
        getattr(L, f"ref_{k}_pk1")(fp, l1[i], l2[i], P[i].ctypes.data_as(dp))
        c = np.zeros(81)
        getattr(L, f"ref_{k}_dpdf")(fp, l1[i], l2[i], c.ctypes.data_as(dp))
        C[i] = c
    np.savez_compressed(os.path.join(HERE, "material_golden.npz"), Fv=Fv,
                        kind=np.array(kinds, dtype=np.uint8), lame1=l1, lame2=l2,
                        psi=psi, P=P, C=C)
    print("wrote", n, "cases")


if __name__ == "__main__":
    main()
