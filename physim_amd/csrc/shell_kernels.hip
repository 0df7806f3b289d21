// shell_kernels.hip -- thin-shell elements for gfx950: StVK membrane
// triangles and discrete-shell hinges.
//
// Both write the same per-element outputs as the tet kernel (12-value
// gradient, 78-value packed Hessian over four connectivity slots), so the
// node gather, slot-map assembly and PCG are shared.  A membrane's fourth
// slot repeats its third node and receives zeros.  One element per lane,
// 256-lane workgroups, FP64 per-workgroup energy partials.
//
// Membrane (MaterialModels.h:376-440, the 2D StVK of FEMMem_StVK_*.mcg):
//   F = R + D (3x2), R = sum_a (X_a - X_0) b_a^T (rest, orthonormal columns),
//   D = sum_a (u_a - u_0) b_a^T (displacements), so
//   E = (R^T D + D^T R + D^T D) / 2 carries no O(1) cancellation;
//   Psi = lambda/2 tr(E)^2 + mu tr(E^2), S = lambda tr(E) I + 2 mu E,
//   g_a = A0 t F S b_a,
//   H_ab = A0 t [(b_a.S b_b) I + lambda u_a u_b^T + mu u_b u_a^T
//                + mu (b_a.b_b) F F^T],  u_a = F b_a.
// Hinge (EnergyElement_DiscreteShells.cpp:56-198): e = x1 - x0,
//   N0 = e x (x2 - x0), N1 = (x3 - x0) x e, m = N0/|N0| + N1/|N1|,
//   theta = 2 atan(e^.(m x N0^) / (m.N0^)), E = kB shape0 (theta - theta0)^2,
//   g = 2 kB shape0 (theta - theta0) dtheta, H = 2 kB shape0 dtheta dtheta^T
//   (the SPD Hessian the reference element assembles), with the closed-form
//   dihedral gradient dtheta/dx2 = |e| N0/|N0|^2, dtheta/dx3 = |e| N1/|N1|^2,
//   dtheta/dx0 = ((x2-x1).e^) N0/|N0|^2 + ((x3-x1).e^) N1/|N1|^2,
//   dtheta/dx1 = -((x2-x0).e^) N0/|N0|^2 - ((x3-x0).e^) N1/|N1|^2.
#include "kernels.hpp"
#include "reduce.hpp"

namespace pgn {

template <class T, int FLAGS>
__global__ __launch_bounds__(TET_BLOCK) void membrane_eval_kernel(
    ShellArgs A, double* __restrict__ blockE, double* __restrict__ elemE,
    T* __restrict__ eg, T* __restrict__ eH) {
  __shared__ double red[TET_BLOCK / 64];
  const int l = blockIdx.x * TET_BLOCK + threadIdx.x;
  double Ee = 0.0;
  if (l < A.count) {
    const long e = A.first + l;
    const long m = A.stride;
    const T* __restrict__ prm = reinterpret_cast<const T*>(A.prm);
    const int4 t = A.elems[e];
    const int nd[3] = {t.x, t.y, t.z};
    T dX[2][3], du[2][3];
#pragma unroll
    for (int a = 1; a < 3; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double X0i = A.X0[3L * nd[0] + i], Xai = A.X0[3L * nd[a] + i];
        const double u0 = A.x[3L * nd[0] + i] - X0i, ua = A.x[3L * nd[a] + i] - Xai;
        dX[a - 1][i] = (T)(Xai - X0i);
        du[a - 1][i] = (T)(ua - u0);
      }
    T b[3][2];
    b[1][0] = prm[0 * m + e];
    b[1][1] = prm[1 * m + e];
    b[2][0] = prm[2 * m + e];
    b[2][1] = prm[3 * m + e];
    b[0][0] = -(b[1][0] + b[2][0]);
    b[0][1] = -(b[1][1] + b[2][1]);
    const T At = prm[9 * m + e];
    const T l1 = prm[10 * m + e];
    const T l2 = prm[11 * m + e];
    T R[3][2], D[3][2], F[3][2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        R[i][j] = dX[0][i] * b[1][j] + dX[1][i] * b[2][j];
        D[i][j] = du[0][i] * b[1][j] + du[1][i] * b[2][j];
        F[i][j] = R[i][j] + D[i][j];
      }
    T Eg[2][2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        T s = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
          s += R[i][j] * D[i][k] + D[i][j] * R[i][k] + D[i][j] * D[i][k];
        Eg[j][k] = T(0.5) * s;
      }
    const T trE = Eg[0][0] + Eg[1][1];
    if (FLAGS & EVAL_E) {
      const T psi = T(0.5) * l1 * trE * trE +
                    l2 * (Eg[0][0] * Eg[0][0] + T(2) * Eg[0][1] * Eg[0][1] + Eg[1][1] * Eg[1][1]);
      Ee = (double)(At * psi);
    }
    T S[2][2];
    S[0][0] = l1 * trE + T(2) * l2 * Eg[0][0];
    S[1][1] = l1 * trE + T(2) * l2 * Eg[1][1];
    S[0][1] = S[1][0] = T(2) * l2 * Eg[0][1];
    if (FLAGS & EVAL_G) {
      T P[3][2];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) P[i][j] = F[i][0] * S[0][j] + F[i][1] * S[1][j];
      T* __restrict__ go = eg + e * 12;
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i) go[3 * a + i] = At * (P[i][0] * b[a][0] + P[i][1] * b[a][1]);
#pragma unroll
      for (int i = 0; i < 3; ++i) go[9 + i] = T(0);
    }
    if (FLAGS & EVAL_H) {
      T u[3][3], G[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i) u[a][i] = F[i][0] * b[a][0] + F[i][1] * b[a][1];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) G[i][k] = F[i][0] * F[k][0] + F[i][1] * F[k][1];
      const T Am = At * l2, Al = At * l1;
      T* __restrict__ h = eH + e * TET_H_PACK;
#pragma unroll
      for (int a = 1; a < 4; ++a)
#pragma unroll
        for (int bb = 0; bb < a; ++bb) {
          T* o = h + 9 * (a * (a - 1) / 2 + bb);
          if (a == 3) {
#pragma unroll
            for (int q = 0; q < 9; ++q) o[q] = T(0);
            continue;
          }
          const T s = At * (b[a][0] * (S[0][0] * b[bb][0] + S[0][1] * b[bb][1]) +
                            b[a][1] * (S[1][0] * b[bb][0] + S[1][1] * b[bb][1]));
          const T beta = Am * (b[a][0] * b[bb][0] + b[a][1] * b[bb][1]);
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 3; ++k)
              o[3 * i + k] = Am * u[bb][i] * u[a][k] + beta * G[i][k] + Al * u[a][i] * u[bb][k] +
                             (i == k ? s : T(0));
        }
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        T* o = h + 54 + 6 * a;
        if (a == 3) {
#pragma unroll
          for (int q = 0; q < 6; ++q) o[q] = T(0);
          continue;
        }
        const T s = At * (b[a][0] * (S[0][0] * b[a][0] + S[0][1] * b[a][1]) +
                          b[a][1] * (S[1][0] * b[a][0] + S[1][1] * b[a][1]));
        const T beta = Am * (b[a][0] * b[a][0] + b[a][1] * b[a][1]);
        const T A2 = Am + Al;
        o[0] = A2 * u[a][0] * u[a][0] + beta * G[0][0] + s;
        o[1] = A2 * u[a][1] * u[a][0] + beta * G[1][0];
        o[2] = A2 * u[a][1] * u[a][1] + beta * G[1][1] + s;
        o[3] = A2 * u[a][2] * u[a][0] + beta * G[2][0];
        o[4] = A2 * u[a][2] * u[a][1] + beta * G[2][1];
        o[5] = A2 * u[a][2] * u[a][2] + beta * G[2][2] + s;
      }
    }
    if ((FLAGS & EVAL_E) && elemE) elemE[e] = Ee;
  }
  if ((FLAGS & EVAL_E) && blockE) {
    const double s = block_sum<TET_BLOCK>(Ee, red);
    if (threadIdx.x == 0) blockE[blockIdx.x] = s;
  }
}

template <class T, int FLAGS>
__global__ __launch_bounds__(TET_BLOCK) void hinge_eval_kernel(
    ShellArgs A, double* __restrict__ blockE, double* __restrict__ elemE,
    T* __restrict__ eg, T* __restrict__ eH) {
  __shared__ double red[TET_BLOCK / 64];
  const int l = blockIdx.x * TET_BLOCK + threadIdx.x;
  double Ee = 0.0;
  if (l < A.count) {
    const long e = A.first + l;
    const long m = A.stride;
    const T* __restrict__ prm = reinterpret_cast<const T*>(A.prm);
    const int4 t = A.elems[e];
    const int nd[4] = {t.x, t.y, t.z, t.w};
    // the dihedral angle in FP64 whatever T is: E and g depend on
    // theta - theta0, which FP32 would resolve only to ~1e-7 rad absolute
    double ed[3], a2d[3], a3d[3];
    T ev[3], a2[3], a3[3], b2[3], b3[3];  // e, x2-x0, x3-x0, x2-x1, x3-x1
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double p0 = A.x[3L * nd[0] + i], p1 = A.x[3L * nd[1] + i];
      const double p2 = A.x[3L * nd[2] + i], p3 = A.x[3L * nd[3] + i];
      ed[i] = p1 - p0;
      a2d[i] = p2 - p0;
      a3d[i] = p3 - p0;
      ev[i] = (T)ed[i];
      a2[i] = (T)a2d[i];
      a3[i] = (T)a3d[i];
      b2[i] = (T)(p2 - p1);
      b3[i] = (T)(p3 - p1);
    }
    const T kB = prm[0 * m + e], sh0 = prm[2 * m + e];
    const double th0 = A.prm64[1 * m + e];  // FP64 whatever T is
    double thd;
    {
      const double M0[3] = {ed[1] * a2d[2] - ed[2] * a2d[1], ed[2] * a2d[0] - ed[0] * a2d[2],
                            ed[0] * a2d[1] - ed[1] * a2d[0]};
      const double M1[3] = {a3d[1] * ed[2] - a3d[2] * ed[1], a3d[2] * ed[0] - a3d[0] * ed[2],
                            a3d[0] * ed[1] - a3d[1] * ed[0]};
      const double s0 = 1.0 / sqrt(M0[0] * M0[0] + M0[1] * M0[1] + M0[2] * M0[2]);
      const double s1 = 1.0 / sqrt(M1[0] * M1[0] + M1[1] * M1[1] + M1[2] * M1[2]);
      double n0[3], mv[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        n0[i] = M0[i] * s0;
        mv[i] = n0[i] + M1[i] * s1;
      }
      const double rm = 1.0 / sqrt(mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2]);
#pragma unroll
      for (int i = 0; i < 3; ++i) mv[i] *= rm;
      const double mn[3] = {mv[1] * n0[2] - mv[2] * n0[1], mv[2] * n0[0] - mv[0] * n0[2],
                            mv[0] * n0[1] - mv[1] * n0[0]};
      const double Ld = sqrt(ed[0] * ed[0] + ed[1] * ed[1] + ed[2] * ed[2]);
      const double num = (ed[0] * mn[0] + ed[1] * mn[1] + ed[2] * mn[2]) / Ld;
      const double den = mv[0] * n0[0] + mv[1] * n0[1] + mv[2] * n0[2];
      thd = 2.0 * atan(num / den);
    }
    const T N0[3] = {ev[1] * a2[2] - ev[2] * a2[1], ev[2] * a2[0] - ev[0] * a2[2],
                     ev[0] * a2[1] - ev[1] * a2[0]};
    const T N1[3] = {a3[1] * ev[2] - a3[2] * ev[1], a3[2] * ev[0] - a3[0] * ev[2],
                     a3[0] * ev[1] - a3[1] * ev[0]};
    const T q0 = N0[0] * N0[0] + N0[1] * N0[1] + N0[2] * N0[2];
    const T q1 = N1[0] * N1[0] + N1[1] * N1[1] + N1[2] * N1[2];
    const T L = sqrt(ev[0] * ev[0] + ev[1] * ev[1] + ev[2] * ev[2]);
    const T iL = T(1) / L;
    const T k = kB * sh0;
    const double dthd = thd - th0;
    const T dth = (T)dthd;
    if (FLAGS & EVAL_E) Ee = (double)k * dthd * dthd;
    if (FLAGS & (EVAL_G | EVAL_H)) {
      const T w20 = (a2[0] * ev[0] + a2[1] * ev[1] + a2[2] * ev[2]) * iL;
      const T w30 = (a3[0] * ev[0] + a3[1] * ev[1] + a3[2] * ev[2]) * iL;
      const T w21 = (b2[0] * ev[0] + b2[1] * ev[1] + b2[2] * ev[2]) * iL;
      const T w31 = (b3[0] * ev[0] + b3[1] * ev[1] + b3[2] * ev[2]) * iL;
      const T iq0 = T(1) / q0, iq1 = T(1) / q1;
      T d[4][3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const T c0 = N0[i] * iq0, c1 = N1[i] * iq1;
        d[0][i] = w21 * c0 + w31 * c1;
        d[1][i] = -w20 * c0 - w30 * c1;
        d[2][i] = L * c0;
        d[3][i] = L * c1;
      }
      if (FLAGS & EVAL_G) {
        const T s = T(2) * k * dth;
        T* __restrict__ go = eg + e * 12;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int i = 0; i < 3; ++i) go[3 * a + i] = s * d[a][i];
      }
      if (FLAGS & EVAL_H) {
        const T s = T(2) * k;
        T* __restrict__ h = eH + e * TET_H_PACK;
#pragma unroll
        for (int a = 1; a < 4; ++a)
#pragma unroll
          for (int bb = 0; bb < a; ++bb) {
            T* o = h + 9 * (a * (a - 1) / 2 + bb);
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
              for (int j = 0; j < 3; ++j) o[3 * i + j] = s * d[a][i] * d[bb][j];
          }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          T* o = h + 54 + 6 * a;
          o[0] = s * d[a][0] * d[a][0];
          o[1] = s * d[a][1] * d[a][0];
          o[2] = s * d[a][1] * d[a][1];
          o[3] = s * d[a][2] * d[a][0];
          o[4] = s * d[a][2] * d[a][1];
          o[5] = s * d[a][2] * d[a][2];
        }
      }
    }
    if ((FLAGS & EVAL_E) && elemE) elemE[e] = Ee;
  }
  if ((FLAGS & EVAL_E) && blockE) {
    const double s = block_sum<TET_BLOCK>(Ee, red);
    if (threadIdx.x == 0) blockE[blockIdx.x] = s;
  }
}

template <class T, int F>
static void launch_s(hipStream_t s, int kind, const ShellArgs& a, double* blockE,
                     double* elemE, void* eg, void* eH) {
  const int grid = (a.count + TET_BLOCK - 1) / TET_BLOCK;
  if (grid == 0) return;
  if (kind == SHELL_MEMBRANE)
    hipLaunchKernelGGL((membrane_eval_kernel<T, F>), dim3(grid), dim3(TET_BLOCK), 0, s, a,
                       blockE, elemE, reinterpret_cast<T*>(eg), reinterpret_cast<T*>(eH));
  else
    hipLaunchKernelGGL((hinge_eval_kernel<T, F>), dim3(grid), dim3(TET_BLOCK), 0, s, a,
                       blockE, elemE, reinterpret_cast<T*>(eg), reinterpret_cast<T*>(eH));
}

template <class T>
static void launch_sflags(hipStream_t s, int kind, int flags, const ShellArgs& a,
                          double* blockE, double* elemE, void* eg, void* eH) {
  switch (flags) {
    case EVAL_E: launch_s<T, EVAL_E>(s, kind, a, blockE, elemE, eg, eH); break;
    case EVAL_E | EVAL_G: launch_s<T, EVAL_E | EVAL_G>(s, kind, a, blockE, elemE, eg, eH); break;
    case EVAL_H: launch_s<T, EVAL_H>(s, kind, a, blockE, elemE, eg, eH); break;
    case EVAL_E | EVAL_G | EVAL_H:
      launch_s<T, EVAL_E | EVAL_G | EVAL_H>(s, kind, a, blockE, elemE, eg, eH);
      break;
    default: break;
  }
}

void launch_shell_eval(hipStream_t s, bool fp32, int kind, int flags, const ShellArgs& a,
                       double* blockE, double* elemE, void* eg, void* eH) {
  if (fp32)
    launch_sflags<float>(s, kind, flags, a, blockE, elemE, eg, eH);
  else
    launch_sflags<double>(s, kind, flags, a, blockE, elemE, eg, eH);
}

}  // namespace pgn
