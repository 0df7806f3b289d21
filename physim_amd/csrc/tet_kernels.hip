// tet_kernels.hip -- Tet4 element energy / gradient / Hessian for gfx950.
//
// One element per lane, workgroups of 64-256 lanes (tiles chosen per model
// by timing).  The per-element arithmetic is
// the reference's (EnergyElement_FEM.cpp:125-222: E = V0 Psi(F),
// g = V0 DFDx^T P, H = V0 DFDx^T dP/dF DFDx with 1-point Tet1 quadrature),
// re-derived so that no 9x12 DFDx or 9x9 dP/dF is ever formed:
//   F = I + D,  D = sum_a du_a b_a^T   (du_a: displacement differences, FP64)
//   g_a = V0 P b_a
//   CoNH: H_ab = V0 [mu (b_a.b_b) I + (mu - lambda lnJ) c_b c_a^T
//                    + lambda c_a c_b^T],  c_a = F^-T b_a
//   StVK: H_ab = V0 [(b_a^T S b_b) I + mu (u_b u_a^T + (b_a.b_b) F F^T)
//                    + lambda u_a u_b^T],  u_a = F b_a
// Working from D = F - I keeps the energy and stress free of the O(1)
// cancellations of I1 - 3 and F - F^-T, so FP32 E and g stay accurate at
// small strains.  The inversion test (Cell_Tetra.cpp:27-36, at Position_X)
// is done in FP64 from the current node positions.
#include <cstdlib>
#include <type_traits>

#include "kernels.hpp"
#include "reduce.hpp"

namespace pgn {

template <class T>
__device__ __forceinline__ T xm_log1p(T y);  // y - log1p(y)

template <>
__device__ __forceinline__ float xm_log1p<float>(float y) {
  if (fabsf(y) < 0.25f) {
    float s = 1.0f / 13.0f;
#pragma unroll
    for (int k = 12; k >= 2; --k) s = 1.0f / (float)k - y * s;
    return y * y * s;
  }
  return y - log1pf(y);
}

template <>
__device__ __forceinline__ double xm_log1p<double>(double y) {
  if (fabs(y) < 0.01) {
    double s = 1.0 / 11.0;
#pragma unroll
    for (int k = 10; k >= 2; --k) s = 1.0 / (double)k - y * s;
    return y * y * s;
  }
  return y - log1p(y);
}

template <class T>
__device__ __forceinline__ T t_log1p(T y);
template <>
__device__ __forceinline__ float t_log1p<float>(float y) {
  return (y > -0.5f) ? log1pf(y) : logf(fabsf(1.0f + y));
}
template <>
__device__ __forceinline__ double t_log1p<double>(double y) {
  return (y > -0.5) ? log1p(y) : log(fabs(1.0 + y));
}

// Copies nv values from LDS to HBM as 16-byte vectors (src and dst 16-byte
// aligned), the remainder value by value.
template <class T, int BS>
__device__ __forceinline__ void copy_out(const T* src, T* dst, int nv) {
  using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
  constexpr int W = 16 / sizeof(T);
  const int nw = nv / W;
  const V* s = reinterpret_cast<const V*>(src);
  V* d = reinterpret_cast<V*>(dst);
  for (int q = threadIdx.x; q < nw; q += BS) d[q] = s[q];
  for (int q = W * nw + threadIdx.x; q < nv; q += BS) dst[q] = src[q];
}

// One element: E (into Ee), g (into gv, and stored unless staged) and the
// packed Hessian (into hrow when staged, else straight to eH).
template <class T, int FLAGS, bool STAGE>
__device__ __forceinline__ void tet_element(const TetArgs& A, long e, T* __restrict__ hrow,
                                            double* __restrict__ elemE, T* __restrict__ eg,
                                            T* __restrict__ eH, double& Ee, T (&gv)[12]) {
  if (e < A.m) {
    const long m = A.stride;
    const T* __restrict__ prm = reinterpret_cast<const T*>(A.prm);
    const int4 t = A.tets[e];
    const int nd[4] = {t.x, t.y, t.z, t.w};
    double x0[3], u0[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      x0[i] = A.x[3L * nd[0] + i];
      u0[i] = x0[i] - A.X0[3L * nd[0] + i];
    }
    double ds[3][3];  // [a-1][i] current edge vectors
    T du[3][3];       // [a-1][i] displacement differences
#pragma unroll
    for (int a = 1; a < 4; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double xa = A.x[3L * nd[a] + i];
        const double ua = xa - A.X0[3L * nd[a] + i];
        ds[a - 1][i] = xa - x0[i];
        du[a - 1][i] = (T)(ua - u0[i]);
      }
    const double detDs =
        ds[0][0] * (ds[1][1] * ds[2][2] - ds[1][2] * ds[2][1]) -
        ds[0][1] * (ds[1][0] * ds[2][2] - ds[1][2] * ds[2][0]) +
        ds[0][2] * (ds[1][0] * ds[2][1] - ds[1][1] * ds[2][0]);
    const bool inverted = detDs < 0.0;

    T b[4][3];
#pragma unroll
    for (int c = 0; c < 9; ++c) b[1 + c / 3][c % 3] = prm[c * m + e];
#pragma unroll
    for (int j = 0; j < 3; ++j) b[0][j] = -(b[1][j] + b[2][j] + b[3][j]);
    const T V0 = prm[9 * m + e];
    const T l1 = prm[10 * m + e];
    const T l2 = prm[11 * m + e];
    const int kind = A.kind[e];

    T D[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        D[i][j] = du[0][i] * b[1][j] + du[1][i] * b[2][j] + du[2][i] * b[3][j];

    T P[3][3];
    if (kind == 0) {
      // ---- compressible Neo-Hookean (MaterialModels.h:235-299) ----
      const T trD = D[0][0] + D[1][1] + D[2][2];
      T trD2 = 0, nD2 = 0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          trD2 += D[i][j] * D[j][i];
          nD2 += D[i][j] * D[i][j];
        }
      const T detD = D[0][0] * (D[1][1] * D[2][2] - D[1][2] * D[2][1]) -
                     D[0][1] * (D[1][0] * D[2][2] - D[1][2] * D[2][0]) +
                     D[0][2] * (D[1][0] * D[2][1] - D[1][1] * D[2][0]);
      const T q = T(0.5) * (trD * trD - trD2) + detD;  // J - 1 - tr D
      const T y = trD + q;                             // J - 1
      const T J = T(1) + y;
      const T lnJ = t_log1p<T>(y);                     // ln|J|
      if (FLAGS & EVAL_E) {
        // mu/2 (I1-3) - mu lnJ + lambda/2 ln^2 J, with I1-3 = 2trD + |D|^2
        // and trD - lnJ = (y - log1p y) - q
        const T psi = T(0.5) * l2 * nD2 + l2 * (xm_log1p<T>(y) - q) +
                      T(0.5) * l1 * lnJ * lnJ;
        Ee = inverted ? HUGE_VAL : (double)(V0 * psi);
      }
      // F^-T = cof(F) / J
      T F[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F[i][j] = D[i][j] + (i == j ? T(1) : T(0));
      const T iJ = T(1) / J;
      T Fit[3][3];
      Fit[0][0] = (F[1][1] * F[2][2] - F[1][2] * F[2][1]) * iJ;
      Fit[0][1] = (F[1][2] * F[2][0] - F[1][0] * F[2][2]) * iJ;
      Fit[0][2] = (F[1][0] * F[2][1] - F[1][1] * F[2][0]) * iJ;
      Fit[1][0] = (F[0][2] * F[2][1] - F[0][1] * F[2][2]) * iJ;
      Fit[1][1] = (F[0][0] * F[2][2] - F[0][2] * F[2][0]) * iJ;
      Fit[1][2] = (F[0][1] * F[2][0] - F[0][0] * F[2][1]) * iJ;
      Fit[2][0] = (F[0][1] * F[1][2] - F[0][2] * F[1][1]) * iJ;
      Fit[2][1] = (F[0][2] * F[1][0] - F[0][0] * F[1][2]) * iJ;
      Fit[2][2] = (F[0][0] * F[1][1] - F[0][1] * F[1][0]) * iJ;
      if (FLAGS & EVAL_G) {
        // P = mu (F - F^-T) + lambda lnJ F^-T, F - F^-T = D + F^-T D^T
        const T ll = l1 * lnJ;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const T FitDt = Fit[i][0] * D[j][0] + Fit[i][1] * D[j][1] +
                            Fit[i][2] * D[j][2];
            P[i][j] = l2 * (D[i][j] + FitDt) + ll * Fit[i][j];
          }
      }
      if (FLAGS & EVAL_H) {
        T c[4][3];
#pragma unroll
        for (int a = 1; a < 4; ++a)
#pragma unroll
          for (int i = 0; i < 3; ++i)
            c[a][i] = Fit[i][0] * b[a][0] + Fit[i][1] * b[a][1] + Fit[i][2] * b[a][2];
#pragma unroll
        for (int i = 0; i < 3; ++i) c[0][i] = -(c[1][i] + c[2][i] + c[3][i]);
        const T A0 = V0 * l2, A1 = V0 * (l2 - l1 * lnJ), A2 = V0 * l1;
        T* h = STAGE ? hrow : eH + e * TET_H_PACK;
#pragma unroll
        for (int a = 1; a < 4; ++a)
#pragma unroll
          for (int bb = 0; bb < a; ++bb) {
            const T beta = A0 * (b[a][0] * b[bb][0] + b[a][1] * b[bb][1] + b[a][2] * b[bb][2]);
            T* o = h + 9 * (a * (a - 1) / 2 + bb);
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
              for (int k = 0; k < 3; ++k)
                o[3 * i + k] = A1 * c[bb][i] * c[a][k] + A2 * c[a][i] * c[bb][k] +
                               (i == k ? beta : T(0));
          }
        const T A12 = A1 + A2;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const T beta = A0 * (b[a][0] * b[a][0] + b[a][1] * b[a][1] + b[a][2] * b[a][2]);
          T* o = h + 54 + 6 * a;
          o[0] = A12 * c[a][0] * c[a][0] + beta;
          o[1] = A12 * c[a][1] * c[a][0];
          o[2] = A12 * c[a][1] * c[a][1] + beta;
          o[3] = A12 * c[a][2] * c[a][0];
          o[4] = A12 * c[a][2] * c[a][1];
          o[5] = A12 * c[a][2] * c[a][2] + beta;
        }
      }
    } else {
      // ---- Saint-Venant Kirchhoff (MaterialModels.h:166-230) ----
      T Eg[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          Eg[i][j] = T(0.5) * (D[i][j] + D[j][i] + D[0][i] * D[0][j] +
                               D[1][i] * D[1][j] + D[2][i] * D[2][j]);
      const T trE = Eg[0][0] + Eg[1][1] + Eg[2][2];
      if (FLAGS & EVAL_E) {
        T nE2 = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) nE2 += Eg[i][j] * Eg[i][j];
        const T psi = T(0.5) * l1 * trE * trE + l2 * nE2;
        Ee = inverted ? HUGE_VAL : (double)(V0 * psi);
      }
      T S[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          S[i][j] = T(2) * l2 * Eg[i][j] + (i == j ? l1 * trE : T(0));
      T F[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F[i][j] = D[i][j] + (i == j ? T(1) : T(0));
      if (FLAGS & EVAL_G) {
        // P = F S = S + D S
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
            P[i][j] = S[i][j] + D[i][0] * S[0][j] + D[i][1] * S[1][j] + D[i][2] * S[2][j];
      }
      if (FLAGS & EVAL_H) {
        T u[4][3], Sb[4][3];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            u[a][i] = F[i][0] * b[a][0] + F[i][1] * b[a][1] + F[i][2] * b[a][2];
            Sb[a][i] = S[i][0] * b[a][0] + S[i][1] * b[a][1] + S[i][2] * b[a][2];
          }
        T G[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int k = 0; k < 3; ++k)
            G[i][k] = F[i][0] * F[k][0] + F[i][1] * F[k][1] + F[i][2] * F[k][2];
        const T A0 = V0 * l2, A2 = V0 * l1;
        T* h = STAGE ? hrow : eH + e * TET_H_PACK;
#pragma unroll
        for (int a = 1; a < 4; ++a)
#pragma unroll
          for (int bb = 0; bb < a; ++bb) {
            const T s = V0 * (b[a][0] * Sb[bb][0] + b[a][1] * Sb[bb][1] + b[a][2] * Sb[bb][2]);
            const T beta = A0 * (b[a][0] * b[bb][0] + b[a][1] * b[bb][1] + b[a][2] * b[bb][2]);
            T* o = h + 9 * (a * (a - 1) / 2 + bb);
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
              for (int k = 0; k < 3; ++k)
                o[3 * i + k] = A0 * u[bb][i] * u[a][k] + beta * G[i][k] +
                               A2 * u[a][i] * u[bb][k] + (i == k ? s : T(0));
          }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const T s = V0 * (b[a][0] * Sb[a][0] + b[a][1] * Sb[a][1] + b[a][2] * Sb[a][2]);
          const T beta = A0 * (b[a][0] * b[a][0] + b[a][1] * b[a][1] + b[a][2] * b[a][2]);
          const T A02 = A0 + A2;
          T* o = h + 54 + 6 * a;
          o[0] = A02 * u[a][0] * u[a][0] + beta * G[0][0] + s;
          o[1] = A02 * u[a][1] * u[a][0] + beta * G[1][0];
          o[2] = A02 * u[a][1] * u[a][1] + beta * G[1][1] + s;
          o[3] = A02 * u[a][2] * u[a][0] + beta * G[2][0];
          o[4] = A02 * u[a][2] * u[a][1] + beta * G[2][1];
          o[5] = A02 * u[a][2] * u[a][2] + beta * G[2][2] + s;
        }
      }
    }
    if (FLAGS & EVAL_G) {
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i)
          gv[3 * a + i] = V0 * (P[i][0] * b[a][0] + P[i][1] * b[a][1] + P[i][2] * b[a][2]);
      if constexpr (!STAGE) {
        T* __restrict__ go = eg + (long)e * 12;
#pragma unroll
        for (int k = 0; k < 12; ++k) go[k] = gv[k];
      }
    }
    if ((FLAGS & EVAL_E) && elemE) elemE[e] = Ee;
  }
}

// BS elements per workgroup pass, EPL passes per workgroup (one element per
// lane per pass: 64 x EPL elements per wavefront).  STAGE: the outputs of a
// pass (packed Hessians, 78 values, then gradients, 12 values) are written to
// LDS one element per lane and copied out as one contiguous stretch each
// with 16-byte stores, instead of every lane storing rows that lie 312 B
// (48 B) apart in HBM; a pass's output offsets, BS x 78 (x 12) values, are
// 16-byte aligned.  blockE receives one FP64 energy partial per 64
// consecutive elements (ceil(m / 64) entries), whatever the tile, so every
// tile gives the same bits.
template <class T, int FLAGS, bool STAGE, int BS, int EPL>
__global__ __launch_bounds__(BS) void tet_eval_kernel(TetArgs A, double* __restrict__ blockE,
                                                      double* __restrict__ elemE,
                                                      T* __restrict__ eg, T* __restrict__ eH) {
  constexpr int SH = (FLAGS & EVAL_H) ? BS * TET_H_PACK : 0;
  constexpr int SG = (FLAGS & EVAL_G) ? BS * 12 : 0;
  constexpr int HS = STAGE ? (SH > SG ? SH : SG) : 1;
  static_assert(HS * sizeof(T) <= 96 * 1024, "staging buffer too large");
  __shared__ __attribute__((aligned(16))) T hs[HS];
#pragma unroll 1
  for (int q = 0; q < EPL; ++q) {
    const long e0 = ((long)blockIdx.x * EPL + q) * BS;
    if (e0 >= A.m) break;  // the same for the whole workgroup
    const long e = e0 + threadIdx.x;
    double Ee = 0.0;
    T gv[12];
    tet_element<T, FLAGS, STAGE>(A, e, hs + (STAGE ? threadIdx.x * TET_H_PACK : 0), elemE, eg,
                                 eH, Ee, gv);
    if constexpr (STAGE) {
      const int ne = (int)((A.m - e0 < BS) ? A.m - e0 : BS);
      if constexpr ((FLAGS & EVAL_H) != 0) {
        __syncthreads();
        copy_out<T, BS>(hs, eH + e0 * TET_H_PACK, ne * TET_H_PACK);
      }
      if constexpr ((FLAGS & EVAL_G) != 0) {
        __syncthreads();  // the Hessian copy has read hs
        if (e < A.m) {
#pragma unroll
          for (int k = 0; k < 12; ++k) hs[threadIdx.x * 12 + k] = gv[k];
        }
        __syncthreads();
        copy_out<T, BS>(hs, eg + e0 * 12, ne * 12);
      }
      __syncthreads();  // the next pass writes hs again
    }
    if ((FLAGS & EVAL_E) && blockE) {
      const double w = wave_sum(Ee);
      const long we = e0 + (threadIdx.x & ~63);
      if ((threadIdx.x & 63) == 0 && we < A.m) blockE[we >> 6] = w;
    }
  }
}

// PGN_TET_STAGE=0 in the environment turns the LDS-staged output stores off
static bool stage_enabled() {
  static const bool on = [] {
    const char* v = std::getenv("PGN_TET_STAGE");
    return !(v && v[0] == '0');
  }();
  return on;
}

template <class T, int F, bool STAGE, int BS, int EPL>
static void launch_k(hipStream_t s, const TetArgs& a, double* blockE, double* elemE, void* eg,
                     void* eH) {
  const long per = (long)BS * EPL;
  const int grid = (int)((a.m + per - 1) / per);
  if (grid == 0) return;
  hipLaunchKernelGGL((tet_eval_kernel<T, F, STAGE, BS, EPL>), dim3(grid), dim3(BS), 0, s, a,
                     blockE, elemE, reinterpret_cast<T*>(eg), reinterpret_cast<T*>(eH));
}

// staging needs BS x 78 values of LDS for Hessians: FP32 only
template <class T, int F>
constexpr bool can_stage() {
  return (F & (EVAL_G | EVAL_H)) != 0 && (sizeof(T) == 4 || (F & EVAL_H) == 0);
}

// Tiles (workgroup size x elements per lane): 0 = 256 x 1 (the default),
// 1 = 128 x 1, 2 = 64 x 1, 3 = 256 x 2, 4 = 128 x 2, 5 = 64 x 2.  Only the
// flag sets of the Newton path (E+g, H, E+g+H) have the tuned tiles.
template <class T, int F>
static void launch_t(hipStream_t s, const TetArgs& a, double* blockE, double* elemE, void* eg,
                     void* eH, int tile) {
  constexpr bool ST = can_stage<T, F>();
  constexpr bool tuned = F == (EVAL_E | EVAL_G) || F == EVAL_H || F == (EVAL_E | EVAL_G | EVAL_H);
  if constexpr (tuned) {
    switch (tile) {
      case 1: launch_k<T, F, ST, 128, 1>(s, a, blockE, elemE, eg, eH); return;
      case 2: launch_k<T, F, ST, 64, 1>(s, a, blockE, elemE, eg, eH); return;
      case 3: launch_k<T, F, ST, 256, 2>(s, a, blockE, elemE, eg, eH); return;
      case 4: launch_k<T, F, ST, 128, 2>(s, a, blockE, elemE, eg, eH); return;
      case 5: launch_k<T, F, ST, 64, 2>(s, a, blockE, elemE, eg, eH); return;
      default: break;
    }
  }
  if constexpr (ST) {
    if (stage_enabled()) {
      launch_k<T, F, true, 256, 1>(s, a, blockE, elemE, eg, eH);
      return;
    }
  }
  launch_k<T, F, false, 256, 1>(s, a, blockE, elemE, eg, eH);
}

template <class T>
static void launch_flags(hipStream_t s, int flags, const TetArgs& a, double* blockE,
                         double* elemE, void* eg, void* eH, int tile) {
  switch (flags) {
    case EVAL_E: launch_t<T, EVAL_E>(s, a, blockE, elemE, eg, eH, tile); break;
    case EVAL_E | EVAL_G: launch_t<T, EVAL_E | EVAL_G>(s, a, blockE, elemE, eg, eH, tile); break;
    case EVAL_H: launch_t<T, EVAL_H>(s, a, blockE, elemE, eg, eH, tile); break;
    case EVAL_E | EVAL_G | EVAL_H:
      launch_t<T, EVAL_E | EVAL_G | EVAL_H>(s, a, blockE, elemE, eg, eH, tile);
      break;
    case EVAL_G: launch_t<T, EVAL_G>(s, a, blockE, elemE, eg, eH, tile); break;
    case EVAL_G | EVAL_H: launch_t<T, EVAL_G | EVAL_H>(s, a, blockE, elemE, eg, eH, tile); break;
    case EVAL_E | EVAL_H: launch_t<T, EVAL_E | EVAL_H>(s, a, blockE, elemE, eg, eH, tile); break;
    default: break;
  }
}

void launch_tet_eval(hipStream_t s, bool fp32, int flags, const TetArgs& a,
                     double* blockE, double* elemE, void* eg, void* eH, int tile) {
  if (fp32)
    launch_flags<float>(s, flags, a, blockE, elemE, eg, eH, tile);
  else
    launch_flags<double>(s, flags, a, blockE, elemE, eg, eH, tile);
}

}  // namespace pgn
