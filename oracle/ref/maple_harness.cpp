// maple_harness.cpp -- C-ABI wrappers around the reference's own Maple element
// kernels, compiled in place from the reference checkout (never copied).
//
// TEST INFRASTRUCTURE ONLY: builds oracle/_ref/libphysim_ref.so, which pins the
// oracle's closed forms to the reference arithmetic.  Each wrapper mirrors the
// reference call site that includes the kernel:
//   FEMVol_CoNH_*   MaterialModels.h:257-297  (MaterialModel_3Din3D_CoNH)
//   FEMVol_StVK_*   MaterialModels.h:187-229  (MaterialModel_3Din3D_StVK)
// The kernels read `Fv` (column-major vec(F)), `lame1`, `lame2` and write the
// energy temporary, `vgx[9]` or `mHx[9][9]`.
#include <math.h>

extern "C" {

double ref_conh_energy(const double* Fv, double lame1, double lame2) {
#include "FEMVol_CoNH_Energy.mcg"
  return t31;
}

void ref_conh_pk1(const double* Fv, double lame1, double lame2, double* out) {
  double vgx[9];
#include "FEMVol_CoNH_Gradient.mcg"
  for (int i = 0; i < 9; ++i) out[i] = vgx[i];
}

void ref_conh_dpdf(const double* Fv, double lame1, double lame2, double* out) {
  double mHx[9][9];
#include "FEMVol_CoNH_Hessian.mcg"
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) out[9 * i + j] = mHx[i][j];
}

double ref_stvk_energy(const double* Fv, double lame1, double lame2) {
#include "FEMVol_StVK_Energy.mcg"
  return t58;
}

void ref_stvk_pk1(const double* Fv, double lame1, double lame2, double* out) {
  double vgx[9];
#include "FEMVol_StVK_Gradient.mcg"
  for (int i = 0; i < 9; ++i) out[i] = vgx[i];
}

void ref_stvk_dpdf(const double* Fv, double lame1, double lame2, double* out) {
  double mHx[9][9];
#include "FEMVol_StVK_Hessian.mcg"
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) out[9 * i + j] = mHx[i][j];
}

}  // extern "C"

// ---- discrete elastic rods ------------------------------------------------
// RodEdge3D*OPT.mcg: EnergyElement_DER_Stretch.cpp:53-120 (reads x12 = edge
// vector, L0, rh, rw, E; writes t15 / vfx[3] / mJx[3][3]).
// getBT*OPT.h: EnergyElement_DER_BendTwist.cpp:67-249.
#include <vector>
#include "getBTEnergyOPT.h"
#include "getBTGradientOPT.h"
#include "getBTHessianOPT.h"
#include "computeReferenceTwistChange.h"

extern "C" {

double ref_rod_stretch_energy(const double* x12, double L0, double rh, double rw,
                              double E) {
#include "RodEdge3DEnergyOPT.mcg"
  return t15;
}

void ref_rod_stretch_force(const double* x12, double L0, double rh, double rw,
                           double E, double* out) {
  std::vector<double> vfx(3);
#include "RodEdge3DForceOPT.mcg"
  for (int i = 0; i < 3; ++i) out[i] = vfx[i];
}

void ref_rod_stretch_jacobian(const double* x12, double L0, double rh, double rw,
                              double E, double* out) {
  double mJx[3][3];
#include "RodEdge3DJacobianOPT.mcg"
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[3 * i + j] = mJx[i][j];
}

// prm: shear[2], young[2], radw[2], radh[2], L0, L1, K0[2], twist0,
//      nRef1[3], tRef1[3], nRef2[3], tRef2[3], rt   (27 values)
double ref_bt_energy(const double* prm, const double* e1, const double* e2,
                     double th1, double th2) {
  double s[2] = {prm[0], prm[1]}, y[2] = {prm[2], prm[3]};
  double w[2] = {prm[4], prm[5]}, h[2] = {prm[6], prm[7]};
  return getBTEnergyOPT(s, y, w, h, prm[8], prm[9], prm + 10, prm[12], prm + 13,
                        prm + 16, prm + 19, prm + 22, prm[25], e1, e2, th1, th2);
}

void ref_bt_gradient(const double* prm, const double* e1, const double* e2,
                     double th1, double th2, double* g8) {
  double s[2] = {prm[0], prm[1]}, y[2] = {prm[2], prm[3]};
  double w[2] = {prm[4], prm[5]}, h[2] = {prm[6], prm[7]};
  getBTGradientOPT(s, y, w, h, prm[8], prm[9], prm + 10, prm[12], prm + 13,
                   prm + 16, prm + 19, prm + 22, prm[25], e1, e2, th1, th2, g8);
}

void ref_bt_hessian(const double* prm, const double* e1, const double* e2,
                    double th1, double th2, double* H64) {
  double s[2] = {prm[0], prm[1]}, y[2] = {prm[2], prm[3]};
  double w[2] = {prm[4], prm[5]}, h[2] = {prm[6], prm[7]};
  getBTHessianOPT(s, y, w, h, prm[8], prm[9], prm + 10, prm[12], prm + 13,
                  prm + 16, prm + 19, prm + 22, prm[25], e1, e2, th1, th2, H64);
}

double ref_reference_twist_change(double* t1, double* n1, double* t2, double* n2,
                                  double rt) {
  return computeReferenceTwistChange(t1, n1, t2, n2, rt);
}

}  // extern "C"

// ---- discrete shells --------------------------------------------------------
// DiscreteShellsBending_*.mcg as called by EnergyElement_DiscreteShells.cpp
// (:56-198): nodes x0, x1 (the hinge edge), x2, x3 (the opposite vertices),
// kB, theta0, shape0.
extern "C" {

double ref_ds_energy(const double* x0, const double* x1, const double* x2,
                     const double* x3, double kB, double theta0, double shape0) {
#include "DiscreteShellsBending_Energy.mcg"
  return t105;
}

void ref_ds_gradient(const double* x0, const double* x1, const double* x2,
                     const double* x3, double kB, double theta0, double shape0,
                     double* out) {
  double vg[12];
#include "DiscreteShellsBending_Gradient.mcg"
  for (int i = 0; i < 12; ++i) out[i] = vg[i];
}

void ref_ds_hessian_spd(const double* x0, const double* x1, const double* x2,
                        const double* x3, double kB, double theta0, double shape0,
                        double* out) {
  (void)theta0;
  double mHspd[12][12];
#include "DiscreteShellsBending_HessianSPD.mcg"
  for (int i = 0; i < 12; ++i)
    for (int j = 0; j < 12; ++j) out[12 * i + j] = mHspd[i][j];
}

double ref_ds_shape(const double* x0, const double* x1, const double* x2,
                    const double* x3) {
#include "DiscreteShellsBending_Shape.mcg"
  return t46;
}

double ref_ds_dihedral(const double* x0, const double* x1, const double* x2,
                       const double* x3) {
#include "DiscreteShellsBending_Dihedral.mcg"
  return t101;
}

// FEMMem_StVK_*.mcg (MaterialModels.h:376-440): Fv = row-major 3x2 F.
double ref_mem_stvk_energy(const double* Fv, double lame1, double lame2) {
#include "FEMMem_StVK_Energy.mcg"
  return t32;
}

void ref_mem_stvk_pk1(const double* Fv, double lame1, double lame2, double* out) {
  double vgx[6];
#include "FEMMem_StVK_Gradient.mcg"
  for (int i = 0; i < 6; ++i) out[i] = vgx[i];
}

void ref_mem_stvk_dpdf(const double* Fv, double lame1, double lame2, double* out) {
  double mHx[6][6];
#include "FEMMem_StVK_Hessian.mcg"
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) out[6 * i + j] = mHx[i][j];
}

}  // extern "C"
