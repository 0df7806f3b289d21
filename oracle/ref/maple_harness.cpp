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
