// rod_kernels.hip -- discrete elastic rods (DER) on gfx950.
//
// Layout: every rod is a chain of vertices; all per-element data is indexed
// by vertex v.  Edge slot v is the edge (v, v+1) (valid unless v is the last
// vertex of its rod), bend-twist slot v is the element centred at v (valid
// for interior vertices).  DoFs are vertex-major [x, y, z, theta_v] where
// theta_v is the material angle of edge v (the last vertex of a rod carries a
// dummy angle that is held fixed), so a rod's Hessian is a symmetric band of
// half-width 10.
//
// Elements (reference: EnergyElement_DER_Stretch.cpp:53-120,
// EnergyElement_DER_BendTwist.cpp:67-249): the stretch energy is closed form;
// the bend-twist energy is evaluated with forward-mode AD numbers (value,
// gradient and Hessian over e1, e2, theta1, theta2) -- see hyperdual.hpp and
// the model statement in rod_bt_energy below.
#include "hyperdual.hpp"
#include "reduce.hpp"
#include "rods.hpp"

namespace pgn {

constexpr int ROD_BS = 128;

// ---- bend-twist energy -------------------------------------------------------
// t_i = e_i/|e_i|; d_i = PT(u_i -> t_i) n_i (frames of the iteration start);
// a_i = Rot(t_i, theta_i) d_i; kb = t1 x t2 / (1 + t1.t2);
// kappa = (kb.(t1 x a1 + t2 x a2), -kb.(a1 + a2));
// phi = angle about t2 from d2 to Rot(t2, rt) PT(t1 -> t2) d1;
// tau = theta2 - theta1 + rt - phi;
// E = kb0 (kappa0 - K0)^2 + kb1 (kappa1 - K1)^2 + kt (tau - tau0)^2.
template <class S>
__device__ __forceinline__ V3<S> pt3(const V3<S>& a, const V3<S>& b, const V3<S>& v) {
  const V3<S> c = cross3(a, b);
  const S d = dot3(a, b);
  using T = decltype(d.v);
  const S k = dot3(c, v) * hd_recip(d + T(1));
  return add3(add3(scale3(v, d), cross3(c, v)), scale3(c, k));
}

template <class S>
__device__ __forceinline__ V3<S> rot3(const V3<S>& t, const S& c, const S& s, const V3<S>& v) {
  const S tv = dot3(t, v);
  using T = decltype(c.v);
  const S omc = c * T(-1) + T(1);
  return add3(add3(scale3(v, c), scale3(cross3(t, v), s)), scale3(t, tv * omc));
}

template <class T, bool WH>
__device__ T rod_bt_energy(const T* xv, const RodBtConst& c, T* g8, T* h36) {
  using S = HD<T, 8, WH>;
  V3<S> e1, e2, n1, u1, n2, u2;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    e1.c[i] = hd_var<T, 8, WH>(xv[i], i);
    e2.c[i] = hd_var<T, 8, WH>(xv[3 + i], 3 + i);
    n1.c[i] = hd_const<T, 8, WH>((T)c.n1[i]);
    u1.c[i] = hd_const<T, 8, WH>((T)c.u1[i]);
    n2.c[i] = hd_const<T, 8, WH>((T)c.n2[i]);
    u2.c[i] = hd_const<T, 8, WH>((T)c.u2[i]);
  }
  const S th1 = hd_var<T, 8, WH>(xv[6], 6), th2 = hd_var<T, 8, WH>(xv[7], 7);
  const V3<S> t1 = scale3(e1, hd_recip(hd_sqrt(dot3(e1, e1))));
  const V3<S> t2 = scale3(e2, hd_recip(hd_sqrt(dot3(e2, e2))));
  const V3<S> d1 = pt3(u1, t1, n1), d2 = pt3(u2, t2, n2);
  const V3<S> a1 = rot3(t1, hd_cos(th1), hd_sin(th1), d1);
  const V3<S> a2 = rot3(t2, hd_cos(th2), hd_sin(th2), d2);
  const V3<S> kb = scale3(cross3(t1, t2), hd_recip(dot3(t1, t2) + T(1)));
  const S k0 = dot3(kb, add3(cross3(t1, a1), cross3(t2, a2)));
  const S k1 = dot3(kb, add3(a1, a2)) * T(-1);
  const V3<S> Q = rot3(t2, hd_const<T, 8, WH>((T)c.crt), hd_const<T, 8, WH>((T)c.srt),
                       pt3(t1, t2, d1));
  const S phi = hd_atan2(dot3(cross3(t2, d2), Q), dot3(d2, Q));
  const S dk0 = k0 + (T)(-c.K0), dk1 = k1 + (T)(-c.K1);
  const S dt = ((th2 - th1) - phi) + (T)(c.rt - c.tau0);
  const S E = dk0 * dk0 * (T)c.kb0 + dk1 * dk1 * (T)c.kb1 + dt * dt * (T)c.kt;
  if (g8)
#pragma unroll
    for (int i = 0; i < 8; ++i) g8[i] = E.g[i];
  if (WH && h36)
#pragma unroll
    for (int i = 0; i < 36; ++i) h36[i] = E.h[i];
  return E.v;
}

__device__ __forceinline__ RodBtConst load_bt_const(const RodArgs& A, long v) {
  RodBtConst c;
  const double* fr = A.frame;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    c.u1[i] = fr[6 * (v - 1) + i];
    c.n1[i] = fr[6 * (v - 1) + 3 + i];
    c.u2[i] = fr[6 * v + i];
    c.n2[i] = fr[6 * v + 3 + i];
  }
  const double* p = A.btc + 6 * v;  // kb0 kb1 kt K0 K1 tau0
  c.kb0 = p[0];
  c.kb1 = p[1];
  c.kt = p[2];
  c.K0 = p[3];
  c.K1 = p[4];
  c.tau0 = p[5];
  c.rt = A.rt[v];
  c.crt = cos(c.rt);
  c.srt = sin(c.rt);
  return c;
}

// ---- element kernels ---------------------------------------------------------
// stretch: E = 1/2 k L0 (|e|/L0 - 1)^2 (k = E pi rh rw); gradient of the edge
// vector (3, head = +, tail = -); Hessian block J (6 values, lower triangle).
template <bool WH>
__global__ __launch_bounds__(ROD_BS) void rod_stretch_kernel(RodArgs A, double* __restrict__ es,
                                                              double* __restrict__ gs,
                                                              double* __restrict__ Js) {
  const long v = blockIdx.x * (long)ROD_BS + threadIdx.x;
  if (v >= A.nv) return;
  if (A.live && !A.live[A.rod_of[v]]) return;
  if (!(A.flags[v] & ROD_HAS_EDGE)) {
    es[v] = 0.0;
    return;
  }
  double e[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) e[i] = A.x[4 * (v + 1) + i] - A.x[4 * v + i];
  const double l = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
  const double L0 = A.L0[v], k = A.kst[v];
  const double s = l / L0 - 1.0;
  es[v] = 0.5 * k * L0 * s * s;
  const double t[3] = {e[0] / l, e[1] / l, e[2] / l};
#pragma unroll
  for (int i = 0; i < 3; ++i) gs[3 * v + i] = k * s * t[i];
  if (WH) {
    const double a = k / L0, b = k * s / l;
    int q = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j, ++q)
        Js[6 * v + q] = a * t[i] * t[j] + b * ((i == j ? 1.0 : 0.0) - t[i] * t[j]);
  }
}

// EG: writes the element energy and gradient (FP64 outputs from arithmetic
// in T); WH: the packed Hessian in T.  The mixed path runs <double, false,
// true> for E and g and <float, true, false> for the Hessian alone.
template <class T, bool WH, bool EG>
__global__ __launch_bounds__(ROD_BS) void rod_bt_kernel(RodArgs A, double* __restrict__ eb,
                                                         double* __restrict__ gb,
                                                         T* __restrict__ Hb) {
  const long v = blockIdx.x * (long)ROD_BS + threadIdx.x;
  if (v >= A.nv) return;
  if (A.live && !A.live[A.rod_of[v]]) return;
  if (!(A.flags[v] & ROD_HAS_BT)) {
    if (EG) eb[v] = 0.0;
    return;
  }
  const RodBtConst c = load_bt_const(A, v);
  T xv[8];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    xv[i] = (T)(A.x[4 * v + i] - A.x[4 * (v - 1) + i]);
    xv[3 + i] = (T)(A.x[4 * (v + 1) + i] - A.x[4 * v + i]);
  }
  xv[6] = (T)A.x[4 * (v - 1) + 3];
  xv[7] = (T)A.x[4 * v + 3];
  T g8[8], h36[WH ? 36 : 1];
  const T E = rod_bt_energy<T, WH>(xv, c, g8, WH ? h36 : nullptr);
  if (EG) {
    eb[v] = (double)E;
#pragma unroll
    for (int i = 0; i < 8; ++i) gb[8 * v + i] = (double)g8[i];
  }
  if (WH)
#pragma unroll
    for (int i = 0; i < 36; ++i) Hb[36 * v + i] = h36[i];
}

// The bend-twist element on caller-given inputs, one element per lane: the
// parameter vector of the reference's getBT* kernels (G, Y, rw, rh and the
// lengths of both edges, K0, K1, tau0, n1, u1, n2, u2, reference twist), the
// edge vectors and the two material angles; E, g8 and the packed Hessian
// (36, WH) in T.
template <class T, bool WH>
__global__ __launch_bounds__(ROD_BS) void rod_bt_raw_kernel(long count,
                                                             const double* __restrict__ prm,
                                                             const double* __restrict__ e1,
                                                             const double* __restrict__ e2,
                                                             const double* __restrict__ th,
                                                             double* __restrict__ E,
                                                             double* __restrict__ g,
                                                             T* __restrict__ H) {
  const long k = blockIdx.x * (long)ROD_BS + threadIdx.x;
  if (k >= count) return;
  const double* p = prm + 27 * k;
  const double pi = 3.14159265358979323846;
  const double G = 0.5 * (p[0] + p[1]), Y = 0.5 * (p[2] + p[3]);
  const double rw = 0.5 * (p[4] + p[5]), rh = 0.5 * (p[6] + p[7]);
  const double lbar = 0.5 * (p[8] + p[9]);
  RodBtConst c;
  c.kb0 = pi / 8.0 * Y * rw * rh * rh * rh / lbar;
  c.kb1 = pi / 8.0 * Y * rw * rw * rw * rh / lbar;
  c.kt = pi / 8.0 * G * rw * rh * (rw * rw + rh * rh) / lbar;
  c.K0 = p[10];
  c.K1 = p[11];
  c.tau0 = p[12];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    c.n1[i] = p[13 + i];
    c.u1[i] = p[16 + i];
    c.n2[i] = p[19 + i];
    c.u2[i] = p[22 + i];
  }
  c.rt = p[25];
  c.crt = cos(c.rt);
  c.srt = sin(c.rt);
  T xv[8];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    xv[i] = (T)e1[3 * k + i];
    xv[3 + i] = (T)e2[3 * k + i];
  }
  xv[6] = (T)th[2 * k];
  xv[7] = (T)th[2 * k + 1];
  T g8[8], h36[WH ? 36 : 1];
  const T v = rod_bt_energy<T, WH>(xv, c, g8, WH ? h36 : nullptr);
  E[k] = (double)v;
#pragma unroll
  for (int i = 0; i < 8; ++i) g[8 * k + i] = (double)g8[i];
  if (WH)
#pragma unroll
    for (int i = 0; i < 36; ++i) H[36 * k + i] = h36[i];
}

// ---- gradient gather + per-vertex energies --------------------------------------
// DoF 4v+c (c < 3) gathers: stretch v-1 (head, +), stretch v (tail, -), and
// the bend-twist elements centred at v+1 (as x_{c-1}: -de1), v (x_c:
// +de1 - de2), v-1 (x_{c+1}: +de2); DoF 4v+3 (theta of edge v) gathers
// bend-twist v (as theta2) and v+1 (as theta1).
__global__ __launch_bounds__(ROD_BS) void rod_gather_kernel(RodArgs A, const double* __restrict__ es,
                                                             const double* __restrict__ gs,
                                                             const double* __restrict__ eb,
                                                             const double* __restrict__ gb,
                                                             double* __restrict__ g,
                                                             double* __restrict__ vE) {
  const long v = blockIdx.x * (long)ROD_BS + threadIdx.x;
  if (v >= A.nv) return;
  if (A.live && !A.live[A.rod_of[v]]) return;
  const unsigned f = A.flags[v];
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  if (f & ROD_HAS_PREV_EDGE)
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] += gs[3 * (v - 1) + i];
  if (f & ROD_HAS_EDGE)
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] -= gs[3 * v + i];
  if ((f & ROD_HAS_EDGE) && (A.flags[v + 1] & ROD_HAS_BT)) {
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] -= gb[8 * (v + 1) + i];
    r[3] += gb[8 * (v + 1) + 6];
  }
  if (f & ROD_HAS_BT) {
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] += gb[8 * v + i] - gb[8 * v + 3 + i];
    r[3] += gb[8 * v + 7];
  }
  if ((f & ROD_HAS_PREV_EDGE) && (A.flags[v - 1] & ROD_HAS_BT))
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] += gb[8 * (v - 1) + 3 + i];
  const double m = A.mass[v];
  double eload = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    r[i] += -m * A.grav[i];
    eload += A.x[4 * v + i] * (-m * A.grav[i]);
  }
  const unsigned fx = A.fixmask[v];
  double gg = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (fx & (1u << i)) r[i] = 0.0;
    g[4 * v + i] = r[i];
    gg += r[i] * r[i];
  }
  vE[3 * v] = es[v] + eb[v];
  vE[3 * v + 1] = eload;
  vE[3 * v + 2] = gg;
}

// one workgroup per rod: fixed-order sums of the per-vertex terms
__global__ __launch_bounds__(256) void rod_reduce_kernel(RodArgs A, const double* __restrict__ vE,
                                                          double* __restrict__ out) {
  __shared__ double red[3 * 256 / 64 + 4];
  const int r = blockIdx.x;
  if (A.live && !A.live[r]) return;
  const long v0 = A.voff[r], v1 = A.voff[r + 1];
  double el = 0.0, ld = 0.0, gg = 0.0;
  for (long v = v0 + threadIdx.x; v < v1; v += 256) {
    el += vE[3 * v];
    ld += vE[3 * v + 1];
    gg += vE[3 * v + 2];
  }
  block_sum2<256>(el, ld, red);
  const double sg = block_sum<256>(gg, red);
  if (threadIdx.x == 0) {
    out[4 * r] = el + ld;
    out[4 * r + 1] = sg;
    out[4 * r + 2] = fabs(el) + fabs(ld);
    out[4 * r + 3] = 0.0;
  }
}

// ---- band Hessian assembly ---------------------------------------------------------
// Entry (d, d-k), k = 0..ROD_BW, of row d.  Element contributions are summed
// in a fixed order: stretch v-1, stretch v, bend-twist v-1, v, v+1.
struct BtMap {  // one local DoF of a bend-twist element as +-variables
  int var0, var1;
  double s0, s1;
};

__device__ __forceinline__ bool bt_local(long d, long c, BtMap& m) {
  // local DoFs of the element centred at vertex c, by global DoF
  const long vd = d >> 2;
  const int comp = (int)(d & 3);
  m.var1 = -1;
  m.s1 = 0.0;
  if (comp < 3) {
    if (vd == c - 1) {
      m.var0 = comp;
      m.s0 = -1.0;
      return true;
    }
    if (vd == c) {
      m.var0 = comp;
      m.s0 = 1.0;
      m.var1 = 3 + comp;
      m.s1 = -1.0;
      return true;
    }
    if (vd == c + 1) {
      m.var0 = 3 + comp;
      m.s0 = 1.0;
      return true;
    }
    return false;
  }
  if (vd == c - 1) {
    m.var0 = 6;
    m.s0 = 1.0;
    return true;
  }
  if (vd == c) {
    m.var0 = 7;
    m.s0 = 1.0;
    return true;
  }
  return false;
}

template <class T>
__device__ __forceinline__ double h36_at(const T* h, int i, int j) {
  return (i >= j) ? (double)h[i * (i + 1) / 2 + j] : (double)h[j * (j + 1) / 2 + i];
}

template <class T>
__global__ __launch_bounds__(ROD_BS) void rod_band_kernel(RodArgs A, const double* __restrict__ Js,
                                                          const T* __restrict__ Hb,
                                                          double* __restrict__ band) {
  const long d = blockIdx.x * (long)ROD_BS + threadIdx.x;
  if (d >= 4 * A.nv) return;
  const long v = d >> 2;
  if (A.live && !A.live[A.rod_of[v]]) return;
  const int comp = (int)(d & 3);
  const long dof0 = 4L * A.voff[A.rod_of[v]];  // first DoF of this rod
  double row[ROD_BW + 1];
#pragma unroll
  for (int k = 0; k <= ROD_BW; ++k) row[k] = 0.0;
  const unsigned f = A.flags[v];
  // stretch blocks: H6 = [[J, -J], [-J, J]] over (x_tail, x_head)
  if (comp < 3) {
    if (f & ROD_HAS_PREV_EDGE) {  // edge v-1: this vertex is the head
      const double* J = Js + 6 * (v - 1);
#pragma unroll
      for (int c2 = 0; c2 < 3; ++c2) {
        const int a = comp > c2 ? comp : c2, b = comp > c2 ? c2 : comp;
        const double j = J[a * (a + 1) / 2 + b];
        row[4 + comp - c2] -= j;                      // (x_v, x_{v-1})
        if (c2 <= comp) row[comp - c2] += j;          // (x_v, x_v)
      }
    }
    if (f & ROD_HAS_EDGE) {  // edge v: this vertex is the tail
      const double* J = Js + 6 * v;
#pragma unroll
      for (int c2 = 0; c2 <= 2; ++c2)
        if (c2 <= comp) row[comp - c2] += J[comp * (comp + 1) / 2 + c2];
    }
  }
  // bend-twist elements centred at v-1, v, v+1
#pragma unroll
  for (int dc = -1; dc <= 1; ++dc) {
    const long c = v + dc;
    if (c < 0 || c >= A.nv) continue;
    if (!(A.flags[c] & ROD_HAS_BT)) continue;
    if (A.rod_of[c] != A.rod_of[v]) continue;
    BtMap mr;
    if (!bt_local(d, c, mr)) continue;
    const T* h = Hb + 36 * c;
#pragma unroll
    for (int k = 0; k <= ROD_BW; ++k) {
      const long dc2 = d - k;
      if (dc2 < dof0) break;
      BtMap mc;
      if (!bt_local(dc2, c, mc)) continue;
      double s = mr.s0 * mc.s0 * h36_at(h, mr.var0, mc.var0);
      if (mc.var1 >= 0) s += mr.s0 * mc.s1 * h36_at(h, mr.var0, mc.var1);
      if (mr.var1 >= 0) {
        s += mr.s1 * mc.s0 * h36_at(h, mr.var1, mc.var0);
        if (mc.var1 >= 0) s += mr.s1 * mc.s1 * h36_at(h, mr.var1, mc.var1);
      }
      row[k] += s;
    }
  }
  // FixMatrixDoFs: fixed rows / columns become identity
  const bool fr = (A.fixmask[v] >> comp) & 1u;
#pragma unroll
  for (int k = 0; k <= ROD_BW; ++k) {
    const long dc2 = d - k;
    double val = row[k];
    if (dc2 < dof0) {
      val = 0.0;
    } else {
      const bool fc = (A.fixmask[dc2 >> 2] >> (dc2 & 3)) & 1u;
      if (fr || fc) val = (k == 0) ? 1.0 : 0.0;
    }
    band[(ROD_BW + 1) * d + k] = val;
  }
}

// ---- banded LDL^T per rod (LinearSolver_EigenLDLT::SolveInternal) -------------------
// One 64-lane workgroup per rod.  Factorization without pivoting of
// band + shift, singular test |D_jj| < sinT, indefiniteness test
// b^.x^ < -regT, accuracy test, and the cumulative regularization loop.
// work: (ROD_BW+1) values per DoF (D on k = 0, L(i, i-k) on k = 1..ROD_BW).
//
// Left-looking factorization, one column j per step: D_j is a reduction over
// the ROD_BW previous columns, and lane t (< ROD_BW) computes L(j+1+t, j) from
// the rows kept in an LDS window of ROD_BW+1 rows (row i in slot i mod W).
// The triangular solves keep the last ROD_BW values in lanes 1..ROD_BW and
// slide them one lane per step.  Norms, the indefiniteness test and the
// accuracy residual are sums over the rod's DoFs by the whole group.
constexpr int LDLT_BS = 64;
// Rows of the band (and of L, b, x in the solves) are staged into LDS RCH
// at a time, chunk c in slot c & 1, so that the sequential steps read LDS
// instead of waiting on a global load each.
constexpr int RCH = 64;
static_assert(RCH > ROD_BW, "a step reads rows j .. j + ROD_BW: two chunks at most");

__device__ __forceinline__ double group_allsum(double v) {
  return wave_sum(v);  // in every lane
}

// Sum of lanes 0..15 (the band lanes), in every lane: an inclusive scan
// within the 16-lane DPP row (row_shr 1, 2, 4, 8; lanes shifted in from
// outside the row read 0) leaves the row total in lane 15, which is then
// read into a scalar register.  DPP moves cost a few cycles where the
// generic shuffle goes through LDS permutes.
template <int N>
__device__ __forceinline__ double row_shr(double v) {
  return __builtin_amdgcn_update_dpp(0.0, v, 0x110 + N, 0xF, 0xF, true);
}
__device__ __forceinline__ double lane15(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), 15);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 15);
  return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));
}
__device__ __forceinline__ double band_allsum(double v) {
  v += row_shr<1>(v);
  v += row_shr<2>(v);
  v += row_shr<4>(v);
  v += row_shr<8>(v);
  return lane15(v);
}

// two such sums at once
__device__ __forceinline__ void band_allsum2(double& a, double& b) {
  a += row_shr<1>(a);
  b += row_shr<1>(b);
  a += row_shr<2>(a);
  b += row_shr<2>(b);
  a += row_shr<4>(a);
  b += row_shr<4>(b);
  a += row_shr<8>(a);
  b += row_shr<8>(b);
  a = lane15(a);
  b = lane15(b);
}

__global__ __launch_bounds__(LDLT_BS) void rod_ldlt_kernel(RodArgs A, const double* __restrict__ band,
                                                          double* __restrict__ work,
                                                          const double* __restrict__ b,
                                                          double* __restrict__ x, int reg_iters,
                                                          double max_error, int* __restrict__ status,
                                                          const int* __restrict__ order,
                                                          const unsigned char* __restrict__ active) {
  constexpr int W = ROD_BW + 1;
  // rings of the last RING rows (RING >= W, a power of two: slot = row & (RING - 1))
  constexpr int RING = 16;
  static_assert(RING >= W && (RING & (RING - 1)) == 0, "ring of rows");
  __shared__ double Lw[RING][W];  // Lw[i & 15][k] = L(i, i - k), k = 1..ROD_BW
  __shared__ double Dw[RING];     // Dw[m & 15] = D(m)
  __shared__ double Bs[2][RCH][W];  // staged rows (band or L) of chunks c, c + 1
  __shared__ double Vs[2][RCH];     // staged b or x of the same chunks
  if ((int)blockIdx.x >= A.nrods) return;
  const int r = order ? order[blockIdx.x] : (int)blockIdx.x;
  if (active && !active[r]) return;
  const int t = threadIdx.x;
  const long d0 = 4L * A.voff[r], d1 = 4L * A.voff[r + 1];
  const long N = d1 - d0;
  // chunk c of a row-major N x W array (rows past N read as zero) into slot
  // c & 1, and of a vector v when given
  auto stage = [&](long c, const double* __restrict__ src, const double* __restrict__ v) {
    if (c < 0) return;
    const long r0 = c * RCH;
    double* dst = &Bs[c & 1][0][0];
    for (int q = t; q < RCH * W; q += LDLT_BS) {
      const long row = r0 + q / W;
      dst[q] = row < N ? src[W * (d0 + row) + q % W] : 0.0;
    }
    if (v)
      for (int q = t; q < RCH; q += LDLT_BS) Vs[c & 1][q] = r0 + q < N ? v[d0 + r0 + q] : 0.0;
  };
  auto brow = [&](long i) -> const double* { return Bs[(i / RCH) & 1][i % RCH]; };
  // |A| over the entry set AMatrixSd stores: lower triangle + the upper
  // entries of every vertex's 3x3 position block
  double nA = 0.0, nbb = 0.0;
  for (long d = d0 + t; d < d1; d += LDLT_BS) {
    const int comp = (int)(d & 3);
    for (int k = 0; k < W; ++k) {
      const double a = band[W * d + k];
      nA += a * a;
      if (comp < 3 && k > 0 && k <= comp) nA += a * a;
    }
    nbb += b[d] * b[d];
  }
  nA = sqrt(group_allsum(nA));
  const double nb = sqrt(group_allsum(nbb));
  double sinT = fmin(1e-6 * nA, 1e-6);
  sinT = fmax(sinT, 1e-9);
  const double regT = 1e-6 * nA;
  int st = 1;
  for (int it = 0; it < reg_iters && st != 0; ++it) {
    double cum = 0.0;  // LinearSolver_EigenLDLT.cpp:68: A += regT 10^(i-1) I, cumulative
    for (int q = 1; q <= it; ++q) cum += regT * pow(10.0, q - 1);
    for (int q = t; q < RING * W; q += LDLT_BS) (&Lw[0][0])[q] = 0.0;
    if (t < RING) Dw[t] = 0.0;
    stage(0, band, b);
    __syncthreads();
    bool bad = false;
    const int Ni = (int)N;  // at most 4 x 500 DoFs
    // the forward substitution L y = b runs inside the factorization: row j
    // of L is complete at step j; lane k in 1..ROD_BW holds y(j - k)
    double win = 0.0;
    for (int j = 0; j < Ni; ++j) {
      if ((j & (RCH - 1)) == 0) {  // rows up to j + ROD_BW: chunk j / RCH + 1 as well
        stage(j / RCH + 1, band, b);
        __syncthreads();
      }
      const long gj = d0 + j;
      const int sj = j & (RING - 1);
      double term = 0.0, ty = 0.0;
      if (t >= 1 && t < W && t <= j) {
        const double l = Lw[sj][t];
        term = l * l * Dw[(j - t) & (RING - 1)];
        ty = l * win;
      }
      band_allsum2(term, ty);  // terms in lanes 1..ROD_BW
      const double dj = brow(j)[0] + cum - term;
      if (!(fabs(dj) >= sinT)) {  // uniform: dj is the same in every lane
        bad = true;
        break;
      }
      const double yj = Vs[(j / RCH) & 1][j & (RCH - 1)] - ty;
      // L(i, j), i = j + 1 + t: (A(i, j) - sum_m L(i, m) L(j, m) D(m)) / D_j over
      // the columns m < j within the band of both rows
      const int i = j + 1 + t;
      const bool row = t < ROD_BW && i < Ni;
      double lij = 0.0;
      if (row) {
        const int o = t + 1;  // i - j
        const int si = i & (RING - 1);
        double sacc = brow(i)[o];
        for (int k = o + 1; k < W && k <= i; ++k)
          sacc -= Lw[si][k] * Lw[sj][k - o] * Dw[(i - k) & (RING - 1)];
        lij = sacc / dj;
      }
      __syncthreads();  // the window of step j has been read
      if (t == 0) {
        Dw[sj] = dj;
        work[W * gj] = dj;
        x[gj] = yj;
      }
      if (row) {
        Lw[i & (RING - 1)][t + 1] = lij;
        work[W * (d0 + i) + t + 1] = lij;
      }
      if (t >= 1 && t < W) Lw[sj][t] = 0.0;  // slot of row j becomes row j + RING
      win = row_shr<1>(win);  // lane k takes lane k - 1 (lanes 1..ROD_BW matter)
      if (t == 1) win = yj;
      __syncthreads();
    }
    if (bad) continue;
    __syncthreads();  // the factor's and y's global rows are complete
    // D z = y
    for (long d = d0 + t; d < d1; d += LDLT_BS) x[d] /= work[W * d];
    __syncthreads();
    // L^T x = z: lane k in 1..ROD_BW holds x(j + k); step j reads rows j ..
    // j + ROD_BW of L (chunks c, c + 1) and z(j)
    win = 0.0;
    if (N > 0) {
      const long cl = (N - 1) / RCH;
      stage(cl + 1, work, nullptr);
      stage(cl, work, x);
      __syncthreads();
    }
    for (int j = Ni - 1; j >= 0; --j) {
      if ((j & (RCH - 1)) == RCH - 1 && j != Ni - 1) {
        __syncthreads();
        stage(j / RCH, work, x);
        __syncthreads();
      }
      const long gj = d0 + j;
      const double term = (t >= 1 && t < W && j + t < Ni) ? brow(j + t)[t] * win : 0.0;
      const double xj = Vs[(j / RCH) & 1][j & (RCH - 1)] - band_allsum(term);
      if (t == 0) x[gj] = xj;
      win = row_shr<1>(win);  // lane k takes lane k - 1 (lanes 1..ROD_BW matter)
      if (t == 1) win = xj;
    }
    __syncthreads();
    // indefiniteness: normalized(b).normalized(x) < -regT
    double nxx = 0.0, bx = 0.0;
    for (long d = d0 + t; d < d1; d += LDLT_BS) {
      nxx += x[d] * x[d];
      bx += b[d] * x[d];
    }
    const double nx = sqrt(group_allsum(nxx));
    bx = group_allsum(bx);
    if (nb > 0.0 && nx > 0.0 && bx / (nb * nx) < -regT) continue;
    // accuracy: |b - (A + shift) x| against maxError (abs and relative)
    double ae = 0.0;
    for (long d = d0 + t; d < d1; d += LDLT_BS) {
      double sv = (band[W * d] + cum) * x[d];
      for (int k = 1; k < W && d - k >= d0; ++k) sv += band[W * d + k] * x[d - k];
      for (int k = 1; k < W && d + k < d1; ++k) sv += band[W * (d + k) + k] * x[d + k];
      ae += (b[d] - sv) * (b[d] - sv);
    }
    ae = sqrt(group_allsum(ae));
    if (ae > max_error && (nb > 0.0 ? ae / nb : ae) > max_error) continue;
    st = 0;
  }
  if (t == 0) status[r] = st;
}

// ---- reference frames and twist (UpdateKinematics) -----------------------------------
// Mesh_DER::AdaptReferenceFrame: frame of edge v parallel-transported from its
// old tangent to the current one (GeometryUtils.cpp:348-367).
__device__ __forceinline__ void pt_d(const double* u, const double* t1, const double* t2,
                                     double* r) {
  const double d = t1[0] * t2[0] + t1[1] * t2[1] + t1[2] * t2[2];
  const double b[3] = {t1[1] * t2[2] - t1[2] * t2[1], t1[2] * t2[0] - t1[0] * t2[2],
                       t1[0] * t2[1] - t1[1] * t2[0]};
  const double bu[3] = {b[1] * u[2] - b[2] * u[1], b[2] * u[0] - b[0] * u[2],
                        b[0] * u[1] - b[1] * u[0]};
  const double k = (b[0] * u[0] + b[1] * u[1] + b[2] * u[2]) / (1.0 + d);
#pragma unroll
  for (int i = 0; i < 3; ++i) r[i] = u[i] * d + b[i] * k + bu[i];
}

__global__ __launch_bounds__(ROD_BS) void rod_frames_kernel(RodArgs A, double* __restrict__ frame) {
  const long v = blockIdx.x * (long)ROD_BS + threadIdx.x;
  if (v >= A.nv || !(A.flags[v] & ROD_HAS_EDGE)) return;
  double t[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) t[i] = A.x[4 * (v + 1) + i] - A.x[4 * v + i];
  const double l = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
#pragma unroll
  for (int i = 0; i < 3; ++i) t[i] /= l;
  double told[3], nold[3], nn[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    told[i] = frame[6 * v + i];
    nold[i] = frame[6 * v + 3 + i];
  }
  pt_d(nold, told, t, nn);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    frame[6 * v + i] = t[i];
    frame[6 * v + 3 + i] = nn[i];
  }
}

// Mesh_DER::ComputeReferenceTwist: rt += -(signed angle about t2 from n2 to
// Rot(t2, rt) PT(t1 -> t2) n1) with the updated frames of edges v-1 and v.
__global__ __launch_bounds__(ROD_BS) void rod_twist_kernel(RodArgs A, const double* __restrict__ frame,
                                                           double* __restrict__ rt) {
  const long v = blockIdx.x * (long)ROD_BS + threadIdx.x;
  if (v >= A.nv || !(A.flags[v] & ROD_HAS_BT)) return;
  const double* t1 = frame + 6 * (v - 1);
  const double* n1 = t1 + 3;
  const double* t2 = frame + 6 * v;
  const double* n2 = t2 + 3;
  double P[3];
  pt_d(n1, t1, t2, P);
  const double r0 = rt[v];
  const double c = cos(r0), s = sin(r0);
  const double tp = t2[0] * P[0] + t2[1] * P[1] + t2[2] * P[2];
  const double txp[3] = {t2[1] * P[2] - t2[2] * P[1], t2[2] * P[0] - t2[0] * P[2],
                         t2[0] * P[1] - t2[1] * P[0]};
  double R[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) R[i] = c * P[i] + s * txp[i] + tp * (1.0 - c) * t2[i];
  const double bn[3] = {t2[1] * n2[2] - t2[2] * n2[1], t2[2] * n2[0] - t2[0] * n2[2],
                        t2[0] * n2[1] - t2[1] * n2[0]};
  const double y = bn[0] * R[0] + bn[1] * R[1] + bn[2] * R[2];
  const double xx = n2[0] * R[0] + n2[1] * R[1] + n2[2] * R[2];
  rt[v] = r0 - atan2(y, xx);
}

// x = xsave + f[rod] * dx for rods in the line search; others untouched
__global__ __launch_bounds__(256) void rod_step_kernel(long nd, const int* __restrict__ rod_of,
                                                       const double* __restrict__ fac,
                                                       const double* __restrict__ xsave,
                                                       const double* __restrict__ dx,
                                                       double* __restrict__ x) {
  const long d = blockIdx.x * 256L + threadIdx.x;
  if (d >= nd) return;
  const double f = fac[rod_of[d >> 2]];
  if (f >= 0.0) x[d] = xsave[d] + f * dx[d];
}

// ---- launchers ------------------------------------------------------------------------
static inline unsigned blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void rod_launch_eval(hipStream_t s, const RodArgs& a, bool hess, bool fp32_hess, RodBufs& b) {
  if (a.nv == 0) return;
  const unsigned g = blocks(a.nv, ROD_BS);
  if (hess && fp32_hess) {
    // FP64 energy and gradient (line search, residual), FP32 Hessian
    hipLaunchKernelGGL(rod_stretch_kernel<true>, dim3(g), dim3(ROD_BS), 0, s, a, b.es, b.gs, b.Js);
    hipLaunchKernelGGL((rod_bt_kernel<double, false, true>), dim3(g), dim3(ROD_BS), 0, s, a, b.eb,
                       b.gb, reinterpret_cast<double*>(b.Hb));
    hipLaunchKernelGGL((rod_bt_kernel<float, true, false>), dim3(g), dim3(ROD_BS), 0, s, a, b.eb,
                       b.gb, reinterpret_cast<float*>(b.Hb));
  } else if (hess) {
    hipLaunchKernelGGL(rod_stretch_kernel<true>, dim3(g), dim3(ROD_BS), 0, s, a, b.es, b.gs, b.Js);
    hipLaunchKernelGGL((rod_bt_kernel<double, true, true>), dim3(g), dim3(ROD_BS), 0, s, a, b.eb,
                       b.gb, reinterpret_cast<double*>(b.Hb));
  } else {
    hipLaunchKernelGGL(rod_stretch_kernel<false>, dim3(g), dim3(ROD_BS), 0, s, a, b.es, b.gs, b.Js);
    hipLaunchKernelGGL((rod_bt_kernel<double, false, true>), dim3(g), dim3(ROD_BS), 0, s, a, b.eb,
                       b.gb, reinterpret_cast<double*>(b.Hb));
  }
  hipLaunchKernelGGL(rod_gather_kernel, dim3(g), dim3(ROD_BS), 0, s, a, b.es, b.gs, b.eb, b.gb,
                     b.g, b.vE);
  hipLaunchKernelGGL(rod_reduce_kernel, dim3(a.nrods), dim3(256), 0, s, a, b.vE, b.rodE);
}

void rod_launch_band(hipStream_t s, const RodArgs& a, bool fp32_hess, RodBufs& b) {
  const unsigned g = blocks(4 * a.nv, ROD_BS);
  if (fp32_hess)
    hipLaunchKernelGGL(rod_band_kernel<float>, dim3(g), dim3(ROD_BS), 0, s, a, b.Js,
                       reinterpret_cast<const float*>(b.Hb), b.band);
  else
    hipLaunchKernelGGL(rod_band_kernel<double>, dim3(g), dim3(ROD_BS), 0, s, a, b.Js,
                       reinterpret_cast<const double*>(b.Hb), b.band);
}

void rod_launch_ldlt(hipStream_t s, const RodArgs& a, RodBufs& b, const double* rhs, double* x,
                     int reg_iters, double max_error, const int* order,
                     const unsigned char* active) {
  if (a.nrods == 0) return;
  hipLaunchKernelGGL(rod_ldlt_kernel, dim3(a.nrods), dim3(LDLT_BS), 0, s, a, b.band, b.work,
                     rhs, x, reg_iters, max_error, b.status, order, active);
}

void rod_launch_bt_raw(hipStream_t s, long count, const double* prm, const double* e1,
                       const double* e2, const double* th, bool fp32, double* E, double* g,
                       void* H) {
  if (count <= 0) return;
  const unsigned gr = blocks(count, ROD_BS);
  if (fp32)
    hipLaunchKernelGGL((rod_bt_raw_kernel<float, true>), dim3(gr), dim3(ROD_BS), 0, s, count, prm,
                       e1, e2, th, E, g, reinterpret_cast<float*>(H));
  else
    hipLaunchKernelGGL((rod_bt_raw_kernel<double, true>), dim3(gr), dim3(ROD_BS), 0, s, count,
                       prm, e1, e2, th, E, g, reinterpret_cast<double*>(H));
}

void rod_launch_kinematics(hipStream_t s, const RodArgs& a, double* frame, double* rt) {
  const unsigned g = blocks(a.nv, ROD_BS);
  hipLaunchKernelGGL(rod_frames_kernel, dim3(g), dim3(ROD_BS), 0, s, a, frame);
  RodArgs a2 = a;
  a2.frame = frame;
  hipLaunchKernelGGL(rod_twist_kernel, dim3(g), dim3(ROD_BS), 0, s, a2, frame, rt);
}

void rod_launch_step(hipStream_t s, const RodArgs& a, const double* fac, const double* xsave,
                     const double* dx, double* x) {
  const long nd = 4 * a.nv;
  hipLaunchKernelGGL(rod_step_kernel, dim3(blocks(nd, 256)), dim3(256), 0, s, nd, a.rod_of,
                     fac, xsave, dx, x);
}

}  // namespace pgn
