/*
 * shells.c -- FP64 CPU restatement of PhySim's thin-shell elements.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *  - hinge bending: EnergyElement_DiscreteShells.cpp:22-198 with the model
 *    its DiscreteShellsBending_*.mcg kernels evaluate.  Nodes x0, x1 span the
 *    hinge edge, x2 / x3 are the vertices opposite it in the two faces
 *    (:30-34).  With e = x1 - x0, n0 = unit(e x (x2 - x0)),
 *    n1 = unit((x3 - x0) x e) and the bisector m = unit(n0 + n1):
 *      theta = 2 atan( (e/|e|).(m x n0) / (m.n0) )      (signed dihedral)
 *      shape = |e|^2 / (4 (|e x (x2-x0)| + |(x3-x0) x e|))
 *      E = kB shape0 (theta - theta0)^2,  g = 2 kB shape0 (theta - theta0) dtheta,
 *      H = 2 kB shape0 dtheta dtheta^T  (the SPD Hessian the element uses, :130)
 *    dtheta is the closed-form dihedral gradient (Bridson et al. 2003):
 *      d/dx2 = |e| N0/|N0|^2, d/dx3 = |e| N1/|N1|^2,
 *      d/dx0 = ((x2-x1).e^) N0/|N0|^2 + ((x3-x1).e^) N1/|N1|^2,
 *      d/dx1 = -((x2-x0).e^) N0/|N0|^2 - ((x3-x0).e^) N1/|N1|^2,
 *    N0 = e x (x2-x0), N1 = (x3-x0) x e.
 *  - membrane: Tri3 with 2D Saint-Venant-Kirchhoff (MaterialModels.h:376-440,
 *    FEMMem_StVK_*.mcg reads F as a row-major 3x2 matrix).  F = Ds Dm^-1
 *    (3x2), Dm = the rest edges in the rest triangle's own orthonormal 2D
 *    frame, E = (F^T F - I)/2, Psi = lambda/2 tr(E)^2 + mu tr(E^2),
 *    energy A0 t Psi.  The reference cannot run this element (its Tri3 FEM
 *    path builds a 3x4 shape matrix), so it is pinned by the closed form and
 *    the Maple material kernel, not by the reference element.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

static double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static void cross3(const double* a, const double* b, double* r) {
  const double x = a[1] * b[2] - a[2] * b[1], y = a[2] * b[0] - a[0] * b[2],
               z = a[0] * b[1] - a[1] * b[0];
  r[0] = x;
  r[1] = y;
  r[2] = z;
}

double orc_hinge_theta(const double x[12]) {
  const double *x0 = x, *x1 = x + 3, *x2 = x + 6, *x3 = x + 9;
  double e[3], a2[3], a3[3], n0[3], n1[3], m[3], mn[3];
  for (int i = 0; i < 3; ++i) {
    e[i] = x1[i] - x0[i];
    a2[i] = x2[i] - x0[i];
    a3[i] = x3[i] - x0[i];
  }
  cross3(e, a2, n0);
  cross3(a3, e, n1);
  const double l0 = sqrt(dot3(n0, n0)), l1 = sqrt(dot3(n1, n1)), le = sqrt(dot3(e, e));
  for (int i = 0; i < 3; ++i) {
    n0[i] /= l0;
    n1[i] /= l1;
    m[i] = n0[i] + n1[i];
  }
  const double lm = sqrt(dot3(m, m));
  for (int i = 0; i < 3; ++i) m[i] /= lm;
  cross3(m, n0, mn);
  return 2.0 * atan(dot3(e, mn) / le / dot3(m, n0));
}

double orc_hinge_shape(const double x[12]) {
  const double *x0 = x, *x1 = x + 3, *x2 = x + 6, *x3 = x + 9;
  double e[3], a2[3], a3[3], N0[3], N1[3];
  for (int i = 0; i < 3; ++i) {
    e[i] = x1[i] - x0[i];
    a2[i] = x2[i] - x0[i];
    a3[i] = x3[i] - x0[i];
  }
  cross3(e, a2, N0);
  cross3(a3, e, N1);
  return dot3(e, e) / (4.0 * sqrt(dot3(N0, N0)) + 4.0 * sqrt(dot3(N1, N1)));
}

void orc_hinge_dtheta(const double x[12], double d[12]) {
  const double *x0 = x, *x1 = x + 3, *x2 = x + 6, *x3 = x + 9;
  double e[3], a2[3], a3[3], N0[3], N1[3];
  for (int i = 0; i < 3; ++i) {
    e[i] = x1[i] - x0[i];
    a2[i] = x2[i] - x0[i];
    a3[i] = x3[i] - x0[i];
  }
  cross3(e, a2, N0);
  cross3(a3, e, N1);
  const double L = sqrt(dot3(e, e));
  const double q0 = 1.0 / dot3(N0, N0), q1 = 1.0 / dot3(N1, N1);
  double w20 = 0, w30 = 0, w21 = 0, w31 = 0; /* (x2-x0).e^, (x3-x0).e^, (x2-x1).e^, ... */
  for (int i = 0; i < 3; ++i) {
    w20 += (x2[i] - x0[i]) * e[i];
    w30 += (x3[i] - x0[i]) * e[i];
    w21 += (x2[i] - x1[i]) * e[i];
    w31 += (x3[i] - x1[i]) * e[i];
  }
  w20 /= L;
  w30 /= L;
  w21 /= L;
  w31 /= L;
  for (int i = 0; i < 3; ++i) {
    d[i] = w21 * N0[i] * q0 + w31 * N1[i] * q1;
    d[3 + i] = -w20 * N0[i] * q0 - w30 * N1[i] * q1;
    d[6 + i] = L * N0[i] * q0;
    d[9 + i] = L * N1[i] * q1;
  }
}

double orc_hinge_egh(const double x[12], double kB, double theta0, double shape0,
                     double* g12, double* H144) {
  const double th = orc_hinge_theta(x);
  const double k = kB * shape0;
  if (g12 || H144) {
    double d[12];
    orc_hinge_dtheta(x, d);
    if (g12)
      for (int i = 0; i < 12; ++i) g12[i] = 2.0 * k * (th - theta0) * d[i];
    if (H144)
      for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) H144[12 * i + j] = 2.0 * k * d[i] * d[j];
  }
  return k * (th - theta0) * (th - theta0);
}

/* rest triangle in its own 2D frame: t1 along edge 01, t2 = n x t1 */
void orc_membrane_rest(const double X[9], double Dminv[4], double* area) {
  double e1[3], e2[3], n[3], t1[3], t2[3];
  for (int i = 0; i < 3; ++i) {
    e1[i] = X[3 + i] - X[i];
    e2[i] = X[6 + i] - X[i];
  }
  cross3(e1, e2, n);
  const double ln = sqrt(dot3(n, n)), l1 = sqrt(dot3(e1, e1));
  for (int i = 0; i < 3; ++i) {
    n[i] /= ln;
    t1[i] = e1[i] / l1;
  }
  cross3(n, t1, t2);
  const double a = dot3(e1, t1), b = dot3(e2, t1), c = dot3(e1, t2), d = dot3(e2, t2);
  const double det = a * d - b * c; /* Dm = [[a, b], [c, d]] */
  Dminv[0] = d / det;
  Dminv[1] = -b / det;
  Dminv[2] = -c / det;
  Dminv[3] = a / det;
  *area = 0.5 * fabs(det);
}

/* Psi of row-major 3x2 Fv (the FEMMem_StVK_*.mcg argument layout) */
double orc_mem_psi(const double* Fv, double lame1, double lame2) {
  const double E00 = 0.5 * (Fv[0] * Fv[0] + Fv[2] * Fv[2] + Fv[4] * Fv[4] - 1.0);
  const double E11 = 0.5 * (Fv[1] * Fv[1] + Fv[3] * Fv[3] + Fv[5] * Fv[5] - 1.0);
  const double E01 = 0.5 * (Fv[0] * Fv[1] + Fv[2] * Fv[3] + Fv[4] * Fv[5]);
  const double tr = E00 + E11;
  return 0.5 * lame1 * tr * tr + lame2 * (E00 * E00 + 2.0 * E01 * E01 + E11 * E11);
}

/* S = lambda tr(E) I + 2 mu E (2x2, row-major) */
static void mem_S(const double* Fv, double lame1, double lame2, double S[4]) {
  const double E00 = 0.5 * (Fv[0] * Fv[0] + Fv[2] * Fv[2] + Fv[4] * Fv[4] - 1.0);
  const double E11 = 0.5 * (Fv[1] * Fv[1] + Fv[3] * Fv[3] + Fv[5] * Fv[5] - 1.0);
  const double E01 = 0.5 * (Fv[0] * Fv[1] + Fv[2] * Fv[3] + Fv[4] * Fv[5]);
  const double tr = E00 + E11;
  S[0] = lame1 * tr + 2.0 * lame2 * E00;
  S[1] = S[2] = 2.0 * lame2 * E01;
  S[3] = lame1 * tr + 2.0 * lame2 * E11;
}

/* P = F S, row-major 3x2 */
void orc_mem_pk1(const double* Fv, double lame1, double lame2, double* P) {
  double S[4];
  mem_S(Fv, lame1, lame2, S);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 2; ++j) P[2 * i + j] = Fv[2 * i] * S[j] + Fv[2 * i + 1] * S[2 + j];
}

/* dP/dF over the row-major 3x2 layout (6x6) */
void orc_mem_dpdf(const double* Fv, double lame1, double lame2, double* C) {
  double S[4];
  mem_S(Fv, lame1, lame2, S);
  /* d P_ij / d F_kl = delta_ik S_lj + F_im dS_mj/dF_kl,
   * dS = lambda tr(dE) I + 2 mu dE, dE_ab = (dF_ka F_kb + F_ka dF_kb)/2 */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 2; ++l) {
          double v = (i == k) ? S[2 * l + j] : 0.0;
          /* dE_ab/dF_kl = (delta_la F_kb + F_ka delta_lb) / 2 */
          for (int mm = 0; mm < 2; ++mm) {
            double dE_mj = 0.5 * (((l == mm) ? Fv[2 * k + j] : 0.0) + ((l == j) ? Fv[2 * k + mm] : 0.0));
            const double trd = Fv[2 * k + l]; /* d tr(E) / dF_kl */
            const double dS = ((mm == j) ? lame1 * trd : 0.0) + 2.0 * lame2 * dE_mj;
            v += Fv[2 * i + mm] * dS;
          }
          C[6 * (2 * i + j) + 2 * k + l] = v;
        }
}

/* element E, g (9: nodes 0..2) and H (9x9) of a membrane triangle */
double orc_membrane_egh(const double x[9], const double Dminv[4], double A0t,
                        double lame1, double lame2, double* g9, double* H81) {
  /* b_a: rows of Dm^-1 for nodes 1, 2; b_0 = -b_1 - b_2 */
  const double b[3][2] = {{-Dminv[0] - Dminv[2], -Dminv[1] - Dminv[3]},
                          {Dminv[0], Dminv[1]},
                          {Dminv[2], Dminv[3]}};
  double Fv[6];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 2; ++j) Fv[2 * i + j] = x[3 + i] * b[1][j] + x[6 + i] * b[2][j] + x[i] * b[0][j];
  const double E = A0t * orc_mem_psi(Fv, lame1, lame2);
  if (g9) {
    double P[6];
    orc_mem_pk1(Fv, lame1, lame2, P);
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < 3; ++i) g9[3 * a + i] = A0t * (P[2 * i] * b[a][0] + P[2 * i + 1] * b[a][1]);
  }
  if (H81) {
    double C[36];
    orc_mem_dpdf(Fv, lame1, lame2, C);
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 3; ++c)
          for (int k = 0; k < 3; ++k) {
            double s = 0.0;
            for (int j = 0; j < 2; ++j)
              for (int l = 0; l < 2; ++l) s += b[a][j] * C[6 * (2 * i + j) + 2 * k + l] * b[c][l];
            H81[9 * (3 * a + i) + 3 * c + k] = A0t * s;
          }
  }
  return E;
}
