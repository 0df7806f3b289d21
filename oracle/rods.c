/*
 * rods.c -- FP64 CPU restatement of PhySim's discrete elastic rods (DER).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *  - stretch element: EnergyElement_DER_Stretch.cpp:53-120
 *      E = 1/2 E pi rh rw L0 (|e|/L0 - 1)^2, e = x_head - x_tail
 *  - bend-twist element: EnergyElement_DER_BendTwist.cpp:25-249 and the
 *    model its getBT*OPT kernels evaluate: reference directors parallel
 *    transported (GeometryUtils.cpp:348-367) from the iteration-start frames
 *    to the current tangents, material directors rotated by the edge angles,
 *    bending strain from the curvature binormal (:285-302), twist
 *    theta_f - theta_e + rt minus the in-step change of the reference twist.
 *    Gradient and Hessian are taken by second-order forward automatic
 *    differentiation (value, gradient and Hessian carried through every
 *    operation), independent of the reference's symbolic code.
 *  - frames: Mesh_DER::AdaptReferenceFrame (Mesh_DER.cpp:318-336),
 *    ComputeReferenceTwist (:337-353), SetRefFrames_TwistFreeRods
 *    (Mesh_Frame.cpp:117-158, ParallalelTransportForward :186-195).
 *  - masses: MassElement_Lumped on edges, m = rho * L / 2 per node
 *    (MassElement_Lumped.cpp:47-54 with VolumeBasis = length).
 *  - linear solve: LinearSolver_EigenLDLT::SolveInternal (LDL^T without
 *    pivoting, singular / indefinite / accuracy checks, cumulative
 *    regularization).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ---- second-order forward AD over 8 variables ------------------------- */
#define NV 8
#define NH 36
typedef struct {
  double v, g[NV], h[NH];
} hd;

static hd hd_c(double c) {
  hd r;
  memset(&r, 0, sizeof(r));
  r.v = c;
  return r;
}
static hd hd_var(double c, int i) {
  hd r = hd_c(c);
  r.g[i] = 1.0;
  return r;
}
static hd hd_add(hd a, hd b) {
  hd r;
  r.v = a.v + b.v;
  for (int i = 0; i < NV; ++i) r.g[i] = a.g[i] + b.g[i];
  for (int i = 0; i < NH; ++i) r.h[i] = a.h[i] + b.h[i];
  return r;
}
static hd hd_sub(hd a, hd b) {
  hd r;
  r.v = a.v - b.v;
  for (int i = 0; i < NV; ++i) r.g[i] = a.g[i] - b.g[i];
  for (int i = 0; i < NH; ++i) r.h[i] = a.h[i] - b.h[i];
  return r;
}
static hd hd_scale(hd a, double s) {
  hd r;
  r.v = a.v * s;
  for (int i = 0; i < NV; ++i) r.g[i] = a.g[i] * s;
  for (int i = 0; i < NH; ++i) r.h[i] = a.h[i] * s;
  return r;
}
static hd hd_addc(hd a, double c) {
  a.v += c;
  return a;
}
static hd hd_mul(hd a, hd b) {
  hd r;
  r.v = a.v * b.v;
  for (int i = 0; i < NV; ++i) r.g[i] = a.v * b.g[i] + b.v * a.g[i];
  int k = 0;
  for (int i = 0; i < NV; ++i)
    for (int j = 0; j <= i; ++j, ++k)
      r.h[k] = a.v * b.h[k] + b.v * a.h[k] + a.g[i] * b.g[j] + a.g[j] * b.g[i];
  return r;
}
/* f(a) with f' = f1, f'' = f2 at a.v */
static hd hd_chain(hd a, double f, double f1, double f2) {
  hd r;
  r.v = f;
  for (int i = 0; i < NV; ++i) r.g[i] = f1 * a.g[i];
  int k = 0;
  for (int i = 0; i < NV; ++i)
    for (int j = 0; j <= i; ++j, ++k) r.h[k] = f1 * a.h[k] + f2 * a.g[i] * a.g[j];
  return r;
}
static hd hd_recip(hd a) {
  const double iv = 1.0 / a.v;
  return hd_chain(a, iv, -iv * iv, 2.0 * iv * iv * iv);
}
static hd hd_div(hd a, hd b) { return hd_mul(a, hd_recip(b)); }
static hd hd_sqrt(hd a) {
  const double s = sqrt(a.v);
  return hd_chain(a, s, 0.5 / s, -0.25 / (a.v * s));
}
static hd hd_sin(hd a) { return hd_chain(a, sin(a.v), cos(a.v), -sin(a.v)); }
static hd hd_cos(hd a) { return hd_chain(a, cos(a.v), -sin(a.v), -cos(a.v)); }
/* atan2(y, x): d = (x dy - y dx) / r2 */
static hd hd_atan2(hd y, hd x) {
  hd r;
  const double r2 = x.v * x.v + y.v * y.v;
  r.v = atan2(y.v, x.v);
  double w[NV];
  for (int i = 0; i < NV; ++i) {
    w[i] = x.v * y.g[i] - y.v * x.g[i];
    r.g[i] = w[i] / r2;
  }
  int k = 0;
  for (int i = 0; i < NV; ++i)
    for (int j = 0; j <= i; ++j, ++k) {
      /* d/dj of w_i / r2, symmetrized */
      const double dwij = x.g[j] * y.g[i] + x.v * y.h[k] - y.g[j] * x.g[i] - y.v * x.h[k];
      const double dwji = x.g[i] * y.g[j] + x.v * y.h[k] - y.g[i] * x.g[j] - y.v * x.h[k];
      const double dr2j = 2.0 * (x.v * x.g[j] + y.v * y.g[j]);
      const double dr2i = 2.0 * (x.v * x.g[i] + y.v * y.g[i]);
      const double hij = dwij / r2 - w[i] * dr2j / (r2 * r2);
      const double hji = dwji / r2 - w[j] * dr2i / (r2 * r2);
      r.h[k] = 0.5 * (hij + hji);
    }
  return r;
}

typedef struct {
  hd c[3];
} hv;

static hv hv_const(const double* a) {
  hv r;
  for (int i = 0; i < 3; ++i) r.c[i] = hd_c(a[i]);
  return r;
}
static hd hv_dot(hv a, hv b) {
  return hd_add(hd_add(hd_mul(a.c[0], b.c[0]), hd_mul(a.c[1], b.c[1])),
                hd_mul(a.c[2], b.c[2]));
}
static hv hv_cross(hv a, hv b) {
  hv r;
  r.c[0] = hd_sub(hd_mul(a.c[1], b.c[2]), hd_mul(a.c[2], b.c[1]));
  r.c[1] = hd_sub(hd_mul(a.c[2], b.c[0]), hd_mul(a.c[0], b.c[2]));
  r.c[2] = hd_sub(hd_mul(a.c[0], b.c[1]), hd_mul(a.c[1], b.c[0]));
  return r;
}
static hv hv_scale(hv a, hd s) {
  hv r;
  for (int i = 0; i < 3; ++i) r.c[i] = hd_mul(a.c[i], s);
  return r;
}
static hv hv_add(hv a, hv b) {
  hv r;
  for (int i = 0; i < 3; ++i) r.c[i] = hd_add(a.c[i], b.c[i]);
  return r;
}
/* PT(a -> b) v = (a.b) v + (a x b) x v + ((a x b).v / (1 + a.b)) (a x b) */
static hv hv_pt(hv a, hv b, hv v) {
  const hv c = hv_cross(a, b);
  const hd d = hv_dot(a, b);
  const hd k = hd_div(hv_dot(c, v), hd_addc(d, 1.0));
  return hv_add(hv_add(hv_scale(v, d), hv_cross(c, v)), hv_scale(c, k));
}
/* Rodrigues rotation of v about unit t: cos v + sin t x v + (t.v)(1-cos) t */
static hv hv_rot(hv t, hd c, hd s, hv v) {
  const hd tv = hv_dot(t, v);
  return hv_add(hv_add(hv_scale(v, c), hv_scale(hv_cross(t, v), s)),
                hv_scale(t, hd_mul(tv, hd_addc(hd_scale(c, -1.0), 1.0))));
}
static hv hv_normalized(hv a) { return hv_scale(a, hd_recip(hd_sqrt(hv_dot(a, a)))); }

double orc_rod_bt(const double* prm, const double* e1v, const double* e2v,
                  double th1v, double th2v, double* g8, double* H64) {
  const double pi = 3.14159265358979323846;
  const double G = 0.5 * (prm[0] + prm[1]), Y = 0.5 * (prm[2] + prm[3]);
  const double rw = 0.5 * (prm[4] + prm[5]), rh = 0.5 * (prm[6] + prm[7]);
  const double lbar = 0.5 * (prm[8] + prm[9]);
  const double kb0 = pi / 8.0 * Y * rw * rh * rh * rh / lbar;
  const double kb1 = pi / 8.0 * Y * rw * rw * rw * rh / lbar;
  const double kt = pi / 8.0 * G * rw * rh * (rw * rw + rh * rh) / lbar;
  hv e1, e2;
  for (int i = 0; i < 3; ++i) {
    e1.c[i] = hd_var(e1v[i], i);
    e2.c[i] = hd_var(e2v[i], 3 + i);
  }
  const hd th1 = hd_var(th1v, 6), th2 = hd_var(th2v, 7);
  const hv n1 = hv_const(prm + 13), u1 = hv_const(prm + 16);
  const hv n2 = hv_const(prm + 19), u2 = hv_const(prm + 22);
  const double rt = prm[25];
  const hv t1 = hv_normalized(e1), t2 = hv_normalized(e2);
  const hv d1 = hv_pt(u1, t1, n1), d2 = hv_pt(u2, t2, n2);
  const hv a1 = hv_rot(t1, hd_cos(th1), hd_sin(th1), d1);
  const hv a2 = hv_rot(t2, hd_cos(th2), hd_sin(th2), d2);
  const hv kb = hv_scale(hv_cross(t1, t2), hd_recip(hd_addc(hv_dot(t1, t2), 1.0)));
  const hd k0 = hv_dot(kb, hv_add(hv_cross(t1, a1), hv_cross(t2, a2)));
  const hd k1 = hd_scale(hv_dot(kb, hv_add(a1, a2)), -1.0);
  const hv Q = hv_rot(t2, hd_c(cos(rt)), hd_c(sin(rt)), hv_pt(t1, t2, d1));
  const hd phi = hd_atan2(hv_dot(hv_cross(t2, d2), Q), hv_dot(d2, Q));
  const hd tau = hd_addc(hd_sub(hd_sub(th2, th1), phi), rt);
  const hd dk0 = hd_addc(k0, -prm[10]), dk1 = hd_addc(k1, -prm[11]);
  const hd dt = hd_addc(tau, -prm[12]);
  const hd E = hd_add(hd_add(hd_scale(hd_mul(dk0, dk0), kb0), hd_scale(hd_mul(dk1, dk1), kb1)),
                      hd_scale(hd_mul(dt, dt), kt));
  if (g8)
    for (int i = 0; i < NV; ++i) g8[i] = E.g[i];
  if (H64) {
    int k = 0;
    for (int i = 0; i < NV; ++i)
      for (int j = 0; j <= i; ++j, ++k) H64[8 * i + j] = H64[8 * j + i] = E.h[k];
  }
  return E.v;
}

/* EnergyElement_DER_Stretch: x12 = x_head - x_tail, k = E pi rh rw.
 * g (6) over [x_tail, x_head], H (36). */
double orc_rod_stretch(const double* x12, double L0, double k, double* g6,
                       double* H36) {
  const double l = sqrt(x12[0] * x12[0] + x12[1] * x12[1] + x12[2] * x12[2]);
  const double s = l / L0 - 1.0;
  const double E = 0.5 * k * L0 * s * s;
  double t[3] = {x12[0] / l, x12[1] / l, x12[2] / l};
  if (g6)
    for (int i = 0; i < 3; ++i) {
      g6[3 + i] = k * s * t[i];
      g6[i] = -g6[3 + i];
    }
  if (H36) {
    double J[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        J[i][j] = k / L0 * t[i] * t[j] + k * s / l * ((i == j ? 1.0 : 0.0) - t[i] * t[j]);
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        const double sg = (a == b) ? 1.0 : -1.0;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) H36[(3 * a + i) * 6 + 3 * b + j] = sg * J[i][j];
      }
  }
  return E;
}

/* ---- plain 3-vector helpers -------------------------------------------- */
static double d3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static void c3(const double* a, const double* b, double* r) {
  const double x = a[1] * b[2] - a[2] * b[1], y = a[2] * b[0] - a[0] * b[2],
               z = a[0] * b[1] - a[1] * b[0];
  r[0] = x;
  r[1] = y;
  r[2] = z;
}
static void n3(double* a) {
  const double l = sqrt(d3(a, a));
  a[0] /= l;
  a[1] /= l;
  a[2] /= l;
}
/* GeometryUtils.cpp:359-367 parallelTransport(u, t1, t2) */
static void pt3(const double* u, const double* t1, const double* t2, double* r) {
  const double d = d3(t1, t2);
  double b[3], bu[3];
  c3(t1, t2, b);
  c3(b, u, bu);
  const double k = d3(b, u) / (1.0 + d);
  for (int i = 0; i < 3; ++i) r[i] = u[i] * d + b[i] * k + bu[i];
}
static void rot3(const double* t, double ang, const double* v, double* r) {
  const double c = cos(ang), s = sin(ang), tv = d3(t, v);
  double tx[3];
  c3(t, v, tx);
  for (int i = 0; i < 3; ++i) r[i] = c * v[i] + s * tx[i] + tv * (1.0 - c) * t[i];
}

/* computeReferenceTwistChange: minus the signed angle, about tan2, from nor2
 * to Rot(tan2, rt) PT(tan1 -> tan2) nor1 */
double orc_reference_twist_change(const double* tan1, const double* nor1,
                                  const double* tan2, const double* nor2,
                                  double rt) {
  double P[3], R[3], bn[3];
  pt3(nor1, tan1, tan2, P);
  rot3(tan2, rt, P, R);
  c3(tan2, nor2, bn);
  return -atan2(d3(bn, R), d3(nor2, R));
}

/* ---- rod model ---------------------------------------------------------- */
struct orc_rod {
  int n;          /* vertices; edges n-1; DoFs 3n + (n-1) */
  int N;
  double *X0, *x; /* x: full DoF vector (positions then edge angles) */
  double *L0, *k_st;           /* per edge */
  double *prm;                 /* per bend-twist element (vertex 1..n-2): 27 */
  double *ftan, *fnor;         /* per edge reference frame (Frame_X) */
  double *rt;                  /* per vertex reference twist (Twist_X) */
  double *mass;                /* per vertex */
  unsigned char* fixed;        /* N */
  double g_ini[3], g_end[3];
  int g_stages, g_cur, has_g;
  double escale;
};

static void rod_bt_inputs(const orc_rod* R, int v, double* prm, double* e1, double* e2,
                          double* th1, double* th2) {
  /* element at interior vertex v: edges v-1 (tail v-1 -> v) and v */
  memcpy(prm, R->prm + 27L * (v - 1), 27 * sizeof(double));
  for (int i = 0; i < 3; ++i) {
    prm[13 + i] = R->fnor[3 * (v - 1) + i];
    prm[16 + i] = R->ftan[3 * (v - 1) + i];
    prm[19 + i] = R->fnor[3 * v + i];
    prm[22 + i] = R->ftan[3 * v + i];
    e1[i] = R->x[3 * v + i] - R->x[3 * (v - 1) + i];
    e2[i] = R->x[3 * (v + 1) + i] - R->x[3 * v + i];
  }
  prm[25] = R->rt[v];
  *th1 = R->x[3 * R->n + v - 1];
  *th2 = R->x[3 * R->n + v];
}

orc_rod* orc_rod_create(int n, const double* X0, const double* frame0_nor,
                        double young, double shear, double rh, double rw,
                        double line_density) {
  orc_rod* R = (orc_rod*)calloc(1, sizeof(orc_rod));
  R->n = n;
  R->N = 3 * n + (n - 1);
  R->X0 = (double*)malloc(sizeof(double) * 3 * n);
  memcpy(R->X0, X0, sizeof(double) * 3 * n);
  R->x = (double*)calloc(R->N, sizeof(double));
  memcpy(R->x, X0, sizeof(double) * 3 * n);
  R->L0 = (double*)malloc(sizeof(double) * (n - 1));
  R->k_st = (double*)malloc(sizeof(double) * (n - 1));
  R->ftan = (double*)malloc(sizeof(double) * 3 * (n - 1));
  R->fnor = (double*)malloc(sizeof(double) * 3 * (n - 1));
  R->rt = (double*)calloc(n, sizeof(double));
  R->mass = (double*)calloc(n, sizeof(double));
  R->fixed = (unsigned char*)calloc(R->N, 1);
  R->prm = (double*)calloc(27L * (n > 2 ? n - 2 : 1), sizeof(double));
  const double pi = 3.14159265358979323846;
  for (int e = 0; e < n - 1; ++e) {
    double t[3];
    for (int i = 0; i < 3; ++i) t[i] = X0[3 * (e + 1) + i] - X0[3 * e + i];
    R->L0[e] = sqrt(d3(t, t));
    R->k_st[e] = young * pi * rh * rw;
    n3(t);
    memcpy(R->ftan + 3 * e, t, sizeof(t));
    /* lumped mass rho L / 2 per node */
    R->mass[e] += line_density * R->L0[e] / 2.0;
    R->mass[e + 1] += line_density * R->L0[e] / 2.0;
  }
  /* first frame: the given normal; then twist-free transport down the rod */
  double nr[3] = {frame0_nor[0], frame0_nor[1], frame0_nor[2]};
  {
    const double* t = R->ftan;
    const double k = d3(nr, t);
    for (int i = 0; i < 3; ++i) nr[i] -= k * t[i];
    n3(nr);
    memcpy(R->fnor, nr, sizeof(nr));
  }
  for (int e = 1; e < n - 1; ++e) pt3(R->fnor + 3 * (e - 1), R->ftan + 3 * (e - 1), R->ftan + 3 * e,
                                      R->fnor + 3 * e);
  /* bend-twist rest data: EnergyElement_DER_BendTwist::Init (:59-65) */
  for (int v = 1; v < n - 1; ++v) {
    double* p = R->prm + 27L * (v - 1);
    p[0] = p[1] = shear;
    p[2] = p[3] = young;
    p[4] = p[5] = rw;
    p[6] = p[7] = rh;
    p[8] = R->L0[v - 1];
    p[9] = R->L0[v];
    /* rest bend strain with Frame_0, angles 0 (ComputeBendStrain :285-302) */
    const double *te = R->ftan + 3 * (v - 1), *tf = R->ftan + 3 * v;
    const double *ne = R->fnor + 3 * (v - 1), *nf = R->fnor + 3 * v;
    double kb[3], be[3], bf[3];
    c3(te, tf, kb);
    const double den = 2.0 / (1.0 + d3(te, tf));
    for (int i = 0; i < 3; ++i) kb[i] *= den;
    c3(te, ne, be);
    c3(tf, nf, bf);
    double sb[3] = {be[0] + bf[0], be[1] + bf[1], be[2] + bf[2]};
    double sn[3] = {ne[0] + nf[0], ne[1] + nf[1], ne[2] + nf[2]};
    p[10] = 0.5 * d3(sb, kb);
    p[11] = -0.5 * d3(sn, kb);
    p[12] = 0.0; /* rest twist: angles 0, reference twist 0 */
  }
  return R;
}

void orc_rod_destroy(orc_rod* R) {
  if (!R) return;
  free(R->X0);
  free(R->x);
  free(R->L0);
  free(R->k_st);
  free(R->ftan);
  free(R->fnor);
  free(R->rt);
  free(R->mass);
  free(R->fixed);
  free(R->prm);
  free(R);
}

int orc_rod_num_dofs(const orc_rod* R) { return R->N; }
void orc_rod_set_fixed(orc_rod* R, const unsigned char* mask) { memcpy(R->fixed, mask, R->N); }
void orc_rod_set_gravity(orc_rod* R, const double* g_ini, const double* g_end, int stages) {
  R->has_g = 1;
  memcpy(R->g_ini, g_ini, 3 * sizeof(double));
  memcpy(R->g_end, g_end, 3 * sizeof(double));
  R->g_stages = stages < 1 ? 1 : stages;
  R->g_cur = 0;
}
int orc_rod_step_load(orc_rod* R) {
  if (R->has_g && R->g_cur < R->g_stages) {
    R->g_cur++;
    return 1;
  }
  return 0;
}
int orc_rod_fully_loaded(const orc_rod* R) { return !R->has_g || R->g_cur == R->g_stages; }
void orc_rod_get_x(const orc_rod* R, double* x) { memcpy(x, R->x, sizeof(double) * R->N); }
void orc_rod_set_x(orc_rod* R, const double* x) { memcpy(R->x, x, sizeof(double) * R->N); }
void orc_rod_get_frames(const orc_rod* R, double* tan, double* nor, double* rt) {
  if (tan) memcpy(tan, R->ftan, sizeof(double) * 3 * (R->n - 1));
  if (nor) memcpy(nor, R->fnor, sizeof(double) * 3 * (R->n - 1));
  if (rt) memcpy(rt, R->rt, sizeof(double) * R->n);
}

static void cur_g(const orc_rod* R, double* g) {
  if (!R->has_g) {
    g[0] = g[1] = g[2] = 0.0;
    return;
  }
  const double a = (double)R->g_cur / R->g_stages;
  for (int i = 0; i < 3; ++i) g[i] = R->g_ini[i] + (R->g_end[i] - R->g_ini[i]) * a;
}

/* energy, gradient (fixed zeroed) and dense Hessian (fixed identity) */
static double rod_eval(orc_rod* R, double* g, double* H) {
  const int n = R->n, N = R->N;
  if (g) memset(g, 0, sizeof(double) * N);
  if (H) memset(H, 0, sizeof(double) * N * N);
  double Eel = 0.0;
  for (int e = 0; e < n - 1; ++e) {
    double x12[3], ge[6], He[36];
    for (int i = 0; i < 3; ++i) x12[i] = R->x[3 * (e + 1) + i] - R->x[3 * e + i];
    Eel += orc_rod_stretch(x12, R->L0[e], R->k_st[e], g ? ge : NULL, H ? He : NULL);
    const int dof[6] = {3 * e, 3 * e + 1, 3 * e + 2, 3 * e + 3, 3 * e + 4, 3 * e + 5};
    if (g)
      for (int a = 0; a < 6; ++a) g[dof[a]] += ge[a];
    if (H)
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) H[(long)dof[a] * N + dof[b]] += He[6 * a + b];
  }
  for (int v = 1; v < n - 1; ++v) {
    double prm[27], e1[3], e2[3], th1, th2, g8[8], H8[64];
    rod_bt_inputs(R, v, prm, e1, e2, &th1, &th2);
    Eel += orc_rod_bt(prm, e1, e2, th1, th2, g ? g8 : NULL, H ? H8 : NULL);
    /* DeDx (EnergyElement_DER_BendTwist.cpp:43-50): e1 = x_v - x_{v-1},
     * e2 = x_{v+1} - x_v, angles of edges v-1 and v */
    double D[8][11];
    memset(D, 0, sizeof(D));
    for (int i = 0; i < 3; ++i) {
      D[i][i] = -1.0;
      D[i][3 + i] = 1.0;
      D[3 + i][3 + i] = -1.0;
      D[3 + i][6 + i] = 1.0;
    }
    D[6][9] = 1.0;
    D[7][10] = 1.0;
    int dof[11];
    for (int i = 0; i < 9; ++i) dof[i] = 3 * (v - 1) + i;
    dof[9] = 3 * n + v - 1;
    dof[10] = 3 * n + v;
    if (g)
      for (int a = 0; a < 11; ++a) {
        double s = 0;
        for (int p = 0; p < 8; ++p) s += D[p][a] * g8[p];
        g[dof[a]] += s;
      }
    if (H) {
      double T[8][11];
      for (int p = 0; p < 8; ++p)
        for (int b = 0; b < 11; ++b) {
          double s = 0;
          for (int q = 0; q < 8; ++q) s += H8[8 * p + q] * D[q][b];
          T[p][b] = s;
        }
      for (int a = 0; a < 11; ++a)
        for (int b = 0; b < 11; ++b) {
          double s = 0;
          for (int p = 0; p < 8; ++p) s += D[p][a] * T[p][b];
          H[(long)dof[a] * N + dof[b]] += s;
        }
    }
  }
  double gv[3], Eload = 0.0;
  cur_g(R, gv);
  for (int v = 0; v < n; ++v)
    for (int i = 0; i < 3; ++i) {
      Eload += R->x[3 * v + i] * (-R->mass[v] * gv[i]);
      if (g) g[3 * v + i] += -R->mass[v] * gv[i];
    }
  R->escale = fabs(Eel) + fabs(Eload);
  for (int k = 0; k < N; ++k)
    if (R->fixed[k]) {
      if (g) g[k] = 0.0;
      if (H) {
        for (int j = 0; j < N; ++j) H[(long)k * N + j] = H[(long)j * N + k] = 0.0;
        H[(long)k * N + k] = 1.0;
      }
    }
  return Eel + Eload;
}

double orc_rod_energy(orc_rod* R) { return rod_eval(R, NULL, NULL); }
void orc_rod_gradient(orc_rod* R, double* g) { rod_eval(R, g, NULL); }
void orc_rod_hessian(orc_rod* R, double* H) { rod_eval(R, NULL, H); }

/* UpdateKinematics: AdaptReferenceFrame for every edge, then
 * ComputeReferenceTwist for every interior vertex (KEleDER.cpp:94-98,216-220) */
void orc_rod_update_kinematics(orc_rod* R) {
  const int n = R->n;
  for (int e = 0; e < n - 1; ++e) {
    double t[3], nn[3];
    for (int i = 0; i < 3; ++i) t[i] = R->x[3 * (e + 1) + i] - R->x[3 * e + i];
    n3(t);
    pt3(R->fnor + 3 * e, R->ftan + 3 * e, t, nn);
    memcpy(R->ftan + 3 * e, t, sizeof(t));
    memcpy(R->fnor + 3 * e, nn, sizeof(nn));
  }
  for (int v = 1; v < n - 1; ++v)
    R->rt[v] += orc_reference_twist_change(R->ftan + 3 * (v - 1), R->fnor + 3 * (v - 1),
                                           R->ftan + 3 * v, R->fnor + 3 * v, R->rt[v]);
}

/* LinearSolver_EigenLDLT::SolveInternal on a dense symmetric matrix (lower
 * part used), LDL^T without pivoting.  Returns 0 on SUCCESS. */
int orc_dense_ldlt_solve(int N, double* A, const double* b, double* x, double maxError,
                         int regIters, int* reg_used) {
  double nA = 0.0;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j <= i; ++j) nA += A[(long)i * N + j] * A[(long)i * N + j];
  return orc_ldlt_solve_norm(N, A, sqrt(nA), b, x, maxError, regIters, reg_used);
}

/* The same with normA given: LinearSolver::GetMatrixData (LinearSolver.cpp:
 * 57-66) takes the Frobenius norm of the matrix as stored, sinT = max(min(
 * 1e-6 normA, 1e-6), 1e-9), regT = 1e-6 normA; the tries of
 * LinearSolver_EigenLDLT.cpp:63-151: cumulative shift regT 10^(i-1) I, a
 * zero or |D_j| < sinT pivot (factorize failure, singular check) or
 * normalized(b).normalized(x) < -regT (indefinite) or |b - A x| above
 * maxError both absolutely and relatively (inexact) moves to the next try;
 * x keeps its input until a factorization succeeds. */
int orc_ldlt_solve_norm(int N, double* A, double nA, const double* b, double* x,
                        double maxError, int regIters, int* reg_used) {
  double sinT = 1e-6 * nA;
  if (sinT > 1e-6) sinT = 1e-6;
  if (sinT < 1e-9) sinT = 1e-9;
  const double regT = 1e-6 * nA;
  double* L = (double*)malloc(sizeof(double) * N * N);
  double* D = (double*)malloc(sizeof(double) * N);
  double* y = (double*)malloc(sizeof(double) * N);
  int status = 1;
  *reg_used = 0;
  for (int it = 0; it < regIters; ++it) {
    if (it != 0) {
      const double s = regT * pow(10.0, it - 1);
      for (int i = 0; i < N; ++i) A[(long)i * N + i] += s;
      *reg_used = it;
    }
    /* factorize */
    int bad = 0;
    for (int j = 0; j < N && !bad; ++j) {
      double dj = A[(long)j * N + j];
      for (int k = 0; k < j; ++k) dj -= L[(long)j * N + k] * L[(long)j * N + k] * D[k];
      D[j] = dj;
      if (!(fabs(dj) >= sinT)) bad = 1; /* singular check (also NaN) */
      for (int i = j + 1; i < N && !bad; ++i) {
        double s = A[(long)i * N + j];
        for (int k = 0; k < j; ++k) s -= L[(long)i * N + k] * L[(long)j * N + k] * D[k];
        L[(long)i * N + j] = s / dj;
      }
    }
    if (bad) continue;
    for (int i = 0; i < N; ++i) {
      double s = b[i];
      for (int k = 0; k < i; ++k) s -= L[(long)i * N + k] * y[k];
      y[i] = s;
    }
    for (int i = 0; i < N; ++i) y[i] /= D[i];
    for (int i = N - 1; i >= 0; --i) {
      double s = y[i];
      for (int k = i + 1; k < N; ++k) s -= L[(long)k * N + i] * x[k];
      x[i] = s;
    }
    /* indefinite check: normalized(b).normalized(x) < -regT */
    double nb = 0, nx = 0, bx = 0;
    for (int i = 0; i < N; ++i) {
      nb += b[i] * b[i];
      nx += x[i] * x[i];
      bx += b[i] * x[i];
    }
    if (nb > 0 && nx > 0 && bx / (sqrt(nb) * sqrt(nx)) < -regT) continue;
    /* accuracy check */
    double ae = 0;
    for (int i = 0; i < N; ++i) {
      double s = 0;
      for (int j = 0; j < N; ++j) {
        const double a = (j <= i) ? A[(long)i * N + j] : A[(long)j * N + i];
        s += a * x[j];
      }
      ae += (b[i] - s) * (b[i] - s);
    }
    ae = sqrt(ae);
    if (ae > maxError && ae / sqrt(nb) > maxError) continue;
    status = 0;
    break;
  }
  free(L);
  free(D);
  free(y);
  return status;
}

double orc_rod_energy_scale(const orc_rod* R) { return R->escale; }

/* ---- Newton over a rod (OptimSolver::SolveFull with EigenLDLT) ---------- */
static double rp_energy(void* p) { return orc_rod_energy((orc_rod*)p); }
static double rp_escale(void* p) { return ((orc_rod*)p)->escale; }
static void rp_gradient(void* p, double* g) { orc_rod_gradient((orc_rod*)p, g); }
static int rp_solve(void* p, const double* b, double* dx, const orc_newton_opts* o) {
  orc_rod* R = (orc_rod*)p;
  double* H = (double*)malloc(sizeof(double) * R->N * R->N);
  orc_rod_hessian(R, H);
  int reg = 0;
  orc_dense_ldlt_solve(R->N, H, b, dx, o->lsTolMaxError, o->lsNumRegTrials, &reg);
  free(H);
  return 1;
}
static int rp_step(void* p) { return orc_rod_step_load((orc_rod*)p); }
static int rp_loaded(void* p) { return orc_rod_fully_loaded((orc_rod*)p); }
static void rp_get(void* p, double* x) { orc_rod_get_x((orc_rod*)p, x); }
static void rp_set(void* p, const double* x) { orc_rod_set_x((orc_rod*)p, x); }
static void rp_finish(void* p) { orc_rod_update_kinematics((orc_rod*)p); }

void orc_rod_newton(orc_rod* R, const orc_newton_opts* o, orc_newton_result* r) {
  orc_problem P = {R,       R->N,      rp_energy, rp_escale, rp_gradient, rp_solve,
                   rp_step, rp_loaded, rp_get,    rp_set,    rp_finish, NULL};
  orc_newton_generic(&P, o, r);
}
