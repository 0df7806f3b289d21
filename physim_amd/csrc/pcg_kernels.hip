// pcg_kernels.hip -- preconditioned CG on the BSR3 Hessian for gfx950.
//
// The algorithm is the textbook Jacobi-PCG that Eigen's ConjugateGradient
// implements (LinearSolver_EigenCG.cpp:89 solveWithGuess; Saad, Iterative
// Methods, Alg. 9.1) with Eigen's stopping rule |r|^2 < max(tol^2 |b|^2,
// DBL_MIN) and its iteration count.  One iteration k is two launches:
//   K1: beta = rz_k / rz_{k-1};  p_k = z_k + beta p_{k-1} (recomputed per
//       column on the fly, written for the own rows into the other p buffer);
//       q = A p_k;  partial p.q.  Also the convergence test of iteration k-1.
//   K2: alpha = rz_k / p.q;  x += alpha p;  r -= alpha q;  z = M^-1 r;
//       partial r.r, r.z
// No device-scope fences or atomics: every producer workgroup writes one
// partial, and each consumer workgroup of the next launch re-sums the
// partials in a fixed order (the same bits in every workgroup; the kernel
// boundary makes them visible).  r.r / r.z partials are double-buffered by the
// parity of the iteration that produced them; init counts as iteration -1.
#include <hip/hip_ext.h>

#include <algorithm>

#include "kernels.hpp"
#include "pgn_internal.hpp"
#include "reduce.hpp"

namespace pgn {

constexpr int K1_BS = 512;   // SpMV workgroup
constexpr int K2_BS = 256;   // vector-update workgroup
constexpr int K2_NPT = 2;    // nodes per thread of one register-staged batch
constexpr int LPR = 8;       // lanes per row of the plain SpMV / residual kernels
constexpr int PCG_BS = 256;
constexpr int ROWS_PER_BLOCK = PCG_BS / LPR;

static inline int round8(int g) { return g >= 8 ? g - g % 8 : g; }
int pcg_grid1(int n, int lpr) {
  const int rows = K1_BS / lpr;
  int g = (n + rows - 1) / rows;
  return round8(g < 1 ? 1 : (g > PCG_GRID_CAP ? PCG_GRID_CAP : g));
}
constexpr int SELL_BS = 256;  // SELL SpMV workgroup: 4 wavefronts, one slice each per pass
void sell_grid(long ns, int& spw, int& grid) {
  constexpr int WAVES = SELL_BS / 64;
  long per = (ns + PCG_MAXP - 1) / PCG_MAXP;
  if (per < WAVES) per = WAVES;
  per = (per + WAVES - 1) / WAVES * WAVES;
  long g = (ns + per - 1) / per;
  if (g < 1) g = 1;
  if (g >= 8) g = (g + 7) / 8 * 8;
  spw = (int)per;
  grid = (int)g;
}

// streaming launches: one register-staged batch of K2_NPT nodes per thread,
// rounded up to whole XCD rounds
int pcg_grid2(int n) {
  int g = (n + K2_BS * K2_NPT - 1) / (K2_BS * K2_NPT);
  if (g < 1) g = 1;
  if (g >= 8) g = (g + 7) / 8 * 8;
  return g > PCG_GRID_CAP ? PCG_GRID_CAP : g;
}
static inline int grid_rows(int n) {
  int g = (n + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK;
  return g < 1 ? 1 : (g > PCG_GRID_CAP ? PCG_GRID_CAP : g);
}

template <int L, class T>
__device__ __forceinline__ T group_sum(T v) {
  if constexpr (L == 16) {
    // a 16-lane group is a DPP row: an inclusive scan, the total in its last lane
    v += dpp_shr<1>(v);
    v += dpp_shr<2>(v);
    v += dpp_shr<4>(v);
    v += dpp_shr<8>(v);
  } else {
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  }
  return v;  // valid in lane L - 1 of the group (with the shuffles: in every lane)
}

template <class T>
__device__ __forceinline__ T group_sum8(T v) {
  return group_sum<8>(v);
}

// XCD-aware blocked partition of [0, n) rows over the grid.  Workgroups are
// handed to the 8 XCDs round-robin by blockIdx.x, and each XCD has its own
// L2: logical block (b % 8) * (G / 8) + b / 8 gives every XCD one contiguous
// eighth of the rows, so its gathers of neighbouring rows' vector entries
// stay inside its own L2 instead of every XCD fetching the whole vector.
// Each logical block then takes one contiguous chunk of rows.
struct RowRange {
  long lo, hi;
};
__device__ __forceinline__ RowRange block_rows(long n) {
  const int G = gridDim.x, b = blockIdx.x;
  const int lb = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
  const long chunk = (n + G - 1) / G;
  RowRange r;
  r.lo = (long)lb * chunk;
  r.hi = r.lo + chunk < n ? r.lo + chunk : n;
  return r;
}

// The three components of node i.  The streaming kernels load every operand
// of a batch of nodes before they store anything: pointer arguments taken
// from a struct carry no no-alias guarantee, so a store between two loads
// would order them and cost one memory round trip per component.
template <class T>
struct N3 {
  T c[3];
};
template <class T>
__device__ __forceinline__ N3<T> ld3(const T* a, long i) {
  N3<T> o;
  o.c[0] = a[3 * i];
  o.c[1] = a[3 * i + 1];
  o.c[2] = a[3 * i + 2];
  return o;
}
template <class T>
__device__ __forceinline__ void st3(T* a, long i, const N3<T>& v) {
  a[3 * i] = v.c[0];
  a[3 * i + 1] = v.c[1];
  a[3 * i + 2] = v.c[2];
}

// preconditioner operands of node i: the 3x3 block inverse (precond 1) or
// the three Jacobi entries
// Block-Jacobi inverses are stored as upper triangles {m00, m01, m02, m11,
// m12, m22} (exactly symmetric); row p of the full block is (m(p,0),
// m(p,1), m(p,2)) with these indices
__device__ __forceinline__ constexpr int sym3(int p, int q) {
  return p <= q ? (p == 0 ? q : (p == 1 ? 2 + q : 5)) : (q == 0 ? p : (q == 1 ? 2 + p : 5));
}
template <class T>
struct PreM {
  T m[6];
};
template <class T>
__device__ __forceinline__ void ld_precond(int precond, const T* dinv, const T* binv, long i,
                                           PreM<T>& M) {
  if (precond == 1) {
#pragma unroll
    for (int e = 0; e < 6; ++e) M.m[e] = binv[6 * i + e];
  } else {
#pragma unroll
    for (int e = 0; e < 3; ++e) M.m[e] = dinv[3 * i + e];
  }
}
template <class T>
__device__ __forceinline__ N3<T> mul_precond(int precond, const PreM<T>& M, const N3<T>& r) {
  N3<T> z;
  if (precond == 1) {
#pragma unroll
    for (int p = 0; p < 3; ++p)
      z.c[p] = M.m[sym3(p, 0)] * r.c[0] + M.m[sym3(p, 1)] * r.c[1] + M.m[sym3(p, 2)] * r.c[2];
  } else {
#pragma unroll
    for (int p = 0; p < 3; ++p) z.c[p] = M.m[p] * r.c[p];
  }
  return z;
}

template <class T>
__device__ __forceinline__ void apply_precond(int precond, const T* __restrict__ dinv,
                                              const T* __restrict__ binv, long i,
                                              const T r[3], T z[3]) {
  if (precond == 1) {
    const T* M = binv + 6 * i;
#pragma unroll
    for (int p = 0; p < 3; ++p)
      z[p] = M[sym3(p, 0)] * r[0] + M[sym3(p, 1)] * r[1] + M[sym3(p, 2)] * r[2];
  } else {
#pragma unroll
    for (int p = 0; p < 3; ++p) z[p] = dinv[3 * i + p] * r[p];
  }
}

// Fixed-order sum of `cnt` partials (stride `stride`) by a whole workgroup;
// the result is returned to every thread.
template <int BS>
__device__ __forceinline__ double all_sum(const double* __restrict__ a, int cnt, int stride,
                                          double* lds) {
  double v = 0.0;
  for (int i = threadIdx.x; i < cnt; i += BS) v += a[(long)i * stride];
  v = block_sum<BS>(v, lds);
  if (threadIdx.x == 0) lds[BS / 64] = v;
  __syncthreads();
  const double r = lds[BS / 64];
  __syncthreads();
  return r;
}

// Three fixed-order sums in one pass (one latency instead of three); each
// result is bitwise what all_sum returns for it.
template <int BS>
__device__ __forceinline__ void all_sum3(const double* __restrict__ a, int ca, int sa,
                                         const double* __restrict__ b, int cb, int sb,
                                         const double* __restrict__ c, int cc, int scs,
                                         double* lds, double& ra, double& rb, double& rc) {
  double va = 0.0, vb = 0.0, vc = 0.0;
  for (int i = threadIdx.x; i < ca; i += BS) va += a[(long)i * sa];
  for (int i = threadIdx.x; i < cb; i += BS) vb += b[(long)i * sb];
  for (int i = threadIdx.x; i < cc; i += BS) vc += c[(long)i * scs];
  va = wave_sum(va);
  vb = wave_sum(vb);
  vc = wave_sum(vc);
  constexpr int NW = BS / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    lds[w] = va;
    lds[NW + w] = vb;
    lds[2 * NW + w] = vc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0, y = 0.0, z = 0.0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      x += lds[i];
      y += lds[NW + i];
      z += lds[2 * NW + i];
    }
    lds[3 * NW] = x;
    lds[3 * NW + 1] = y;
    lds[3 * NW + 2] = z;
  }
  __syncthreads();
  ra = lds[3 * NW];
  rb = lds[3 * NW + 1];
  rc = lds[3 * NW + 2];
  __syncthreads();
}

// the absolute iteration of a launch (relative to its chunk in a graph)
template <class Args>
__device__ __forceinline__ int chunk_k(const Args& A, int k) {
  return A.kctl ? A.kctl[0] + k : k;
}

__device__ __forceinline__ const double* rr_slot(const PcgArgs& A, int k) {
  return A.part2 + (long)(k & 1) * 2 * PCG_MAXP;  // iteration k's {r.r, r.z}
}
__device__ __forceinline__ int rr_count(const PcgArgs& A, int k) {
  return k < 0 ? A.g1 : A.g2;  // init (k = -1) ran on the K1 grid
}

// ---- init: r = b - A x, z = M^-1 r, p0 = p1 = 0 ------------------------------
template <class T, int L>
__global__ __launch_bounds__(K1_BS) void pcg_init_kernel(PcgArgs A, const T* __restrict__ b,
                                                         int x_zero) {
  __shared__ double red[3 * K1_BS / 64];
  const T* __restrict__ bval = reinterpret_cast<const T*>(A.bval);
  T* __restrict__ x = reinterpret_cast<T*>(A.x);
  T* __restrict__ r = reinterpret_cast<T*>(A.r);
  T* __restrict__ z = reinterpret_cast<T*>(A.z);
  T* __restrict__ p0 = reinterpret_cast<T*>(A.p0);
  T* __restrict__ p1 = reinterpret_cast<T*>(A.p1);
  T* __restrict__ qv = reinterpret_cast<T*>(A.q);
  const int lane = threadIdx.x & (L - 1);
  double bb = 0.0, rr = 0.0, rz = 0.0;
  const RowRange rg = block_rows(A.n);
  for (long row = rg.lo + threadIdx.x / L; row < rg.hi; row += K1_BS / L) {
    T y0 = 0, y1 = 0, y2 = 0;
    if (!x_zero) {
      const long e1 = A.rowptr[row + 1];
      for (long s = A.rowptr[row] + lane; s < e1; s += L) {
        const long j = A.bcol[s];
        const T* B = bval + s * 9;
        const T xj0 = x[3 * j], xj1 = x[3 * j + 1], xj2 = x[3 * j + 2];
        y0 += B[0] * xj0 + B[1] * xj1 + B[2] * xj2;
        y1 += B[3] * xj0 + B[4] * xj1 + B[5] * xj2;
        y2 += B[6] * xj0 + B[7] * xj1 + B[8] * xj2;
      }
      y0 = group_sum<L>(y0);
      y1 = group_sum<L>(y1);
      y2 = group_sum<L>(y2);
    }
    if (lane == L - 1) {
      const long i = row;
      T ri[3] = {b[3 * i] - y0, b[3 * i + 1] - y1, b[3 * i + 2] - y2};
      T zi[3];
      apply_precond<T>(A.precond, reinterpret_cast<const T*>(A.dinv),
                       reinterpret_cast<const T*>(A.binv), i, ri, zi);
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        r[3 * i + p] = ri[p];
        z[3 * i + p] = zi[p];
        p0[3 * i + p] = T(0);
        p1[3 * i + p] = T(0);
        qv[3 * i + p] = T(0);
        if (x_zero) x[3 * i + p] = T(0);
        bb += (double)b[3 * i + p] * (double)b[3 * i + p];
        rr += (double)ri[p] * (double)ri[p];
        rz += (double)ri[p] * (double)zi[p];
      }
    }
  }
  const double sbb = block_sum<K1_BS>(bb, red);
  block_sum2<K1_BS>(rr, rz, red);
  if (threadIdx.x == 0) {
    A.partb[blockIdx.x] = sbb;
    double* o = A.part2 + 2 * PCG_MAXP;  // parity slot of iteration -1
    o[2 * blockIdx.x] = rr;
    o[2 * blockIdx.x + 1] = rz;
    if (blockIdx.x == 0) {
      A.sc->done = 0;
      A.sc->iters = 0;
    }
  }
}

// ---- K1: convergence test of k-1; p_k = z + beta p_{k-1}; q = A p_k; p.q ----
// convergence test of iteration k-1 and beta_k, shared by the fused K1 and
// the split K1a; returns false when the solve is done (all threads agree)
template <int BS>
__device__ __forceinline__ bool pcg_beta(const PcgArgs& A, int k, double* red, double& beta) {
  PcgScalars* sc = A.sc;
  const double* prev = rr_slot(A, k - 1);
  const int cprev = rr_count(A, k - 1);
  double rr, rz, rzm = 1.0;
  if (k == 0) {
    double bb;
    all_sum3<BS>(prev, cprev, 2, prev + 1, cprev, 2, A.partb, A.g1, 1, red, rr, rz, bb);
    double thr = A.tol * A.tol * bb;
    if (thr < 2.2250738585072014e-308) thr = 2.2250738585072014e-308;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->thr = thr;
      sc->bb = bb;
      sc->rr = rr;
    }
    if (bb == 0.0 || rr < thr) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->done = 1;
        sc->iters = 0;
      }
      return false;
    }
    beta = 0.0;
    return true;
  }
  all_sum3<BS>(prev, cprev, 2, prev + 1, cprev, 2, rr_slot(A, k - 2) + 1, rr_count(A, k - 2), 2,
               red, rr, rz, rzm);
  if (rr < sc->thr) {  // iteration k-1 converged (Eigen breaks with i = k-1)
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->done = 1;
      sc->iters = k - 1;
      sc->rr = rr;
    }
    return false;
  }
  beta = rz / rzm;
  return true;
}

template <class T, int L>
__global__ __launch_bounds__(K1_BS) void pcg_k1_kernel(PcgArgs A, int k) {
  k = chunk_k(A, k);
  __shared__ double red[3 * K1_BS / 64 + 4];
  PcgScalars* sc = A.sc;
  if (sc->done) return;
  double beta;
  if (!pcg_beta<K1_BS>(A, k, red, beta)) return;
  const T tb = (T)beta;
  const T* __restrict__ bval = reinterpret_cast<const T*>(A.bval);
  const T* __restrict__ z = reinterpret_cast<const T*>(A.z);
  T* __restrict__ pnew = reinterpret_cast<T*>((k & 1) ? A.p1 : A.p0);
  const T* __restrict__ pold = reinterpret_cast<const T*>((k & 1) ? A.p0 : A.p1);
  T* __restrict__ q = reinterpret_cast<T*>(A.q);
  const int lane = threadIdx.x & (L - 1);
  double pq = 0.0;
  const RowRange rg = block_rows(A.n);
  for (long row = rg.lo + threadIdx.x / L; row < rg.hi; row += K1_BS / L) {
    T y0 = 0, y1 = 0, y2 = 0;
    const long e1 = A.rowptr[row + 1];
    for (long s = A.rowptr[row] + lane; s < e1; s += L) {
      const long j = A.bcol[s];
      const T* B = bval + s * 9;
      const T pj0 = z[3 * j] + tb * pold[3 * j];
      const T pj1 = z[3 * j + 1] + tb * pold[3 * j + 1];
      const T pj2 = z[3 * j + 2] + tb * pold[3 * j + 2];
      y0 += B[0] * pj0 + B[1] * pj1 + B[2] * pj2;
      y1 += B[3] * pj0 + B[4] * pj1 + B[5] * pj2;
      y2 += B[6] * pj0 + B[7] * pj1 + B[8] * pj2;
    }
    y0 = group_sum<L>(y0);
    y1 = group_sum<L>(y1);
    y2 = group_sum<L>(y2);
    if (lane == L - 1) {
      const long i = row;
      const T pi0 = z[3 * i] + tb * pold[3 * i];
      const T pi1 = z[3 * i + 1] + tb * pold[3 * i + 1];
      const T pi2 = z[3 * i + 2] + tb * pold[3 * i + 2];
      pnew[3 * i] = pi0;
      pnew[3 * i + 1] = pi1;
      pnew[3 * i + 2] = pi2;
      q[3 * i] = y0;
      q[3 * i + 1] = y1;
      q[3 * i + 2] = y2;
      pq += (double)pi0 * (double)y0 + (double)pi1 * (double)y1 + (double)pi2 * (double)y2;
    }
  }
  const double s = block_sum<K1_BS>(pq, red);
  if (threadIdx.x == 0) A.part1[blockIdx.x] = s;
}

// ---- split K1: K1a (beta, p = z + beta p_old on own rows, streaming) and K1b
// (plain SpMV q = A p, p.q).  The plain SpMV streams the matrix at copy
// bandwidth; folding the p update into its column gathers costs more than
// the extra launch.
template <class T>
__global__ __launch_bounds__(K2_BS) void pcg_k1a_kernel(PcgArgs A, int k) {
  k = chunk_k(A, k);
  __shared__ double red[3 * K2_BS / 64 + 4];
  if (A.sc->done) return;
  const T* z = reinterpret_cast<const T*>(A.z);
  T* pnew = reinterpret_cast<T*>((k & 1) ? A.p1 : A.p0);
  const T* pold = reinterpret_cast<const T*>((k & 1) ? A.p0 : A.p1);
  const RowRange rg = block_rows(A.n);
  // the first batch is loaded before the partial sums: the two latencies overlap
  N3<T> Z[K2_NPT], P[K2_NPT];
  long i0 = rg.lo + threadIdx.x;
#pragma unroll
  for (int j = 0; j < K2_NPT; ++j) {
    const long i = i0 + (long)j * K2_BS;
    if (i < rg.hi) {
      Z[j] = ld3(z, i);
      P[j] = ld3(pold, i);
    }
  }
  double beta;
  if (!pcg_beta<K2_BS>(A, k, red, beta)) return;
  const T tb = (T)beta;
  for (;;) {
#pragma unroll
    for (int j = 0; j < K2_NPT; ++j) {
      const long i = i0 + (long)j * K2_BS;
      if (i < rg.hi) {
        N3<T> v;
#pragma unroll
        for (int p = 0; p < 3; ++p) v.c[p] = Z[j].c[p] + tb * P[j].c[p];
        st3(pnew, i, v);
      }
    }
    i0 += (long)K2_NPT * K2_BS;
    if (i0 >= rg.hi) break;
#pragma unroll
    for (int j = 0; j < K2_NPT; ++j) {
      const long i = i0 + (long)j * K2_BS;
      if (i < rg.hi) {
        Z[j] = ld3(z, i);
        P[j] = ld3(pold, i);
      }
    }
  }
}

template <class T, int L>
__global__ __launch_bounds__(K1_BS) void pcg_k1b_kernel(PcgArgs A, int k) {
  k = chunk_k(A, k);
  __shared__ double red[K1_BS / 64 + 2];
  if (A.sc->done) return;
  const T* __restrict__ bval = reinterpret_cast<const T*>(A.bval);
  const T* __restrict__ p = reinterpret_cast<const T*>((k & 1) ? A.p1 : A.p0);
  T* __restrict__ q = reinterpret_cast<T*>(A.q);
  const int lane = threadIdx.x & (L - 1);
  double pq = 0.0;
  const RowRange rg = block_rows(A.n);
  for (long row = rg.lo + threadIdx.x / L; row < rg.hi; row += K1_BS / L) {
    T y0 = 0, y1 = 0, y2 = 0;
    const long e1 = A.rowptr[row + 1];
    for (long s = A.rowptr[row] + lane; s < e1; s += L) {
      const long j = A.bcol[s];
      const T* B = bval + s * 9;
      const T pj0 = p[3 * j], pj1 = p[3 * j + 1], pj2 = p[3 * j + 2];
      y0 += B[0] * pj0 + B[1] * pj1 + B[2] * pj2;
      y1 += B[3] * pj0 + B[4] * pj1 + B[5] * pj2;
      y2 += B[6] * pj0 + B[7] * pj1 + B[8] * pj2;
    }
    y0 = group_sum<L>(y0);
    y1 = group_sum<L>(y1);
    y2 = group_sum<L>(y2);
    if (lane == L - 1) {
      const long i = row;
      q[3 * i] = y0;
      q[3 * i + 1] = y1;
      q[3 * i + 2] = y2;
      pq += (double)p[3 * i] * (double)y0 + (double)p[3 * i + 1] * (double)y1 +
            (double)p[3 * i + 2] * (double)y2;
    }
  }
  const double s = block_sum<K1_BS>(pq, red);
  if (threadIdx.x == 0) A.part1[blockIdx.x] = s;
}

// ---- Chronopoulos-Gear PCG (cg1) ------------------------------------------------
// (A. T. Chronopoulos, C. W. Gear, J. Comput. Appl. Math. 25 (1989)): with
// w = A z, gamma = r.z and delta = w.z from the same iteration,
//   beta_k = gamma_k / gamma_{k-1},
//   alpha_k = gamma_k / (delta_k - beta_k gamma_k / alpha_{k-1}),
//   p = z + beta p, q = w + beta q (= A p), x += alpha p, r -= alpha q, z = M^-1 r.
// The SpMV reads z, which the previous launch completed, so each iteration is
// one SpMV launch (w, delta partials) and one streaming launch (all vector
// updates, gamma and r.r partials): no separate p update.  Vectors: z, p0 = p,
// p1 = w, q = q.  Scalars alpha, gamma pass between iterations through cgs
// (double-buffered by parity; block 0 writes them).
template <class T, int L>
__global__ __launch_bounds__(K1_BS) void cg1_a_kernel(PcgArgs A, int k) {
  k = chunk_k(A, k);
  __shared__ double red[3 * K1_BS / 64 + 4];
  PcgScalars* sc = A.sc;
  if (sc->done) return;
  const double* prev = rr_slot(A, k - 1);
  const int cprev = rr_count(A, k - 1);
  double rr, bb, unused;
  all_sum3<K1_BS>(prev, cprev, 2, A.partb, k == 0 ? A.g1 : 0, 1, A.partb, 0, 1, red, rr, bb,
                  unused);
  if (k == 0) {
    double thr = A.tol * A.tol * bb;
    if (thr < 2.2250738585072014e-308) thr = 2.2250738585072014e-308;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->thr = thr;
      sc->bb = bb;
      sc->rr = rr;
    }
    if (bb == 0.0 || rr < thr) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->done = 1;
        sc->iters = 0;
      }
      return;
    }
  } else if (rr < sc->thr) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->done = 1;
      sc->iters = k - 1;
      sc->rr = rr;
    }
    return;
  }
  const T* __restrict__ bval = reinterpret_cast<const T*>(A.bval);
  const T* __restrict__ z = reinterpret_cast<const T*>(A.z);
  T* __restrict__ w = reinterpret_cast<T*>(A.p1);
  const int lane = threadIdx.x & (L - 1);
  double dl = 0.0;
  const RowRange rg = block_rows(A.n);
  for (long row = rg.lo + threadIdx.x / L; row < rg.hi; row += K1_BS / L) {
    T y0 = 0, y1 = 0, y2 = 0;
    const long e1 = A.rowptr[row + 1];
    for (long s = A.rowptr[row] + lane; s < e1; s += L) {
      const long j = A.bcol[s];
      const T* B = bval + s * 9;
      const T zj0 = z[3 * j], zj1 = z[3 * j + 1], zj2 = z[3 * j + 2];
      y0 += B[0] * zj0 + B[1] * zj1 + B[2] * zj2;
      y1 += B[3] * zj0 + B[4] * zj1 + B[5] * zj2;
      y2 += B[6] * zj0 + B[7] * zj1 + B[8] * zj2;
    }
    y0 = group_sum<L>(y0);
    y1 = group_sum<L>(y1);
    y2 = group_sum<L>(y2);
    if (lane == L - 1) {
      const long i = row;
      w[3 * i] = y0;
      w[3 * i + 1] = y1;
      w[3 * i + 2] = y2;
      dl += (double)z[3 * i] * (double)y0 + (double)z[3 * i + 1] * (double)y1 +
            (double)z[3 * i + 2] * (double)y2;
    }
  }
  const double s = block_sum<K1_BS>(dl, red);
  if (threadIdx.x == 0) A.part1[blockIdx.x] = s;
}

// The same step with the SpMV over the SELL-64 copy: a wavefront takes a
// slice of 64 rows, one row per lane, walking the row's blocks in column
// order; every value and column load of the wavefront is 64 consecutive
// entries.  Workgroups take spw consecutive slices (blocked per XCD).
// C16: block columns stored as 16-bit offsets from the slot's row.
template <class T, bool C16>
__global__ __launch_bounds__(SELL_BS) void cg1_sell_kernel(PcgArgs A, int k) {
  k = chunk_k(A, k);
  __shared__ double red[3 * SELL_BS / 64 + 4];
  constexpr int WAVES = SELL_BS / 64;
  PcgScalars* sc = A.sc;
  const T* __restrict__ sval = reinterpret_cast<const T*>(A.sell_val);
  const int* __restrict__ scol = A.sell_col;
  const short* __restrict__ s16 = A.sell_c16;
  const T* __restrict__ z = reinterpret_cast<const T*>(A.z);
  const int G = gridDim.x, b = blockIdx.x;
  const int lb = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // the first block of this wavefront's first slice, loaded before the
  // partial sums (unconditional loads at clamped indexes: every slice has at
  // least one block per lane slot, every row its diagonal block)
  const long s0 = min((long)lb * A.sell_spw + wave, A.sell_ns - 1);
  const long base0 = A.sell_so[s0];
  const int len0 = A.sell_len[min(s0 * 64 + lane, (long)A.n - 1)];
  const long c0 = C16 ? s0 * 64 + lane + s16[base0 + lane] : scol[base0 + lane];
  T v0[9], z0[3];
#pragma unroll
  for (int q = 0; q < 9; ++q) v0[q] = sval[base0 * 9 + lane + 64 * q];
#pragma unroll
  for (int q = 0; q < 3; ++q) z0[q] = z[3 * c0 + q];
  // the flag is read beside the loads above and the partial sums (all in
  // bounds whether or not the solve has ended): their latencies overlap
  const int done = sc->done;
  const double cdot = A.coarse ? *A.cg.cdot : 0.0;  // r.z of the coarse correction
  const double* prev = rr_slot(A, k - 1);
  const int cprev = rr_count(A, k - 1);
  double rr, bb, gam;
  all_sum3<SELL_BS>(prev, cprev, 2, A.partb, k == 0 ? A.g1 : 0, 1, prev + 1, cprev, 2, red, rr,
                    bb, gam);
  if (done) return;
  gam += cdot;
  if (k == 0) {
    double thr = A.tol * A.tol * bb;
    if (thr < 2.2250738585072014e-308) thr = 2.2250738585072014e-308;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->thr = thr;
      sc->bb = bb;
      sc->rr = rr;
    }
    if (bb == 0.0 || rr < thr) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->done = 1;
        sc->iters = 0;
      }
      return;
    }
  } else if (rr < sc->thr) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->done = 1;
      sc->iters = k - 1;
      sc->rr = rr;
    }
    return;
  }
  // beta of this iteration (Chronopoulos-Gear: r.z over the previous r.z)
  const double beta = k == 0 ? 0.0 : gam / A.cgs[2 * ((k - 1) & 1) + 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    A.cgs[4] = beta;
    A.cgs[5] = gam;
  }
  const T tb = (T)beta;
  T* __restrict__ q = reinterpret_cast<T*>(A.q);
  double dl = 0.0;
  for (int t = wave; t < A.sell_spw; t += WAVES) {
    const long slice = (long)lb * A.sell_spw + t;
    if (slice >= A.sell_ns) break;
    const long row = slice * 64 + lane;
    const bool first = t == wave;
    const int len = row < A.n ? (first ? len0 : A.sell_len[row]) : 0;
    const long base = first ? base0 : A.sell_so[slice];
    T y0 = 0, y1 = 0, y2 = 0;
    int j = 0;
    if (first && len > 0) {
      y0 = v0[0] * z0[0] + v0[1] * z0[1] + v0[2] * z0[2];
      y1 = v0[3] * z0[0] + v0[4] * z0[1] + v0[5] * z0[2];
      y2 = v0[6] * z0[0] + v0[7] * z0[1] + v0[8] * z0[2];
      j = 1;
    }
    for (; j < len; ++j) {
      const long c = C16 ? row + s16[base + 64L * j + lane] : scol[base + 64L * j + lane];
      const T* V = sval + (base + 64L * j) * 9 + lane;
      const T a0 = z[3 * c], a1 = z[3 * c + 1], a2 = z[3 * c + 2];
      y0 += V[0] * a0 + V[64] * a1 + V[128] * a2;
      y1 += V[192] * a0 + V[256] * a1 + V[320] * a2;
      y2 += V[384] * a0 + V[448] * a1 + V[512] * a2;
    }
    if (row < A.n) {
      const T q0 = q[3 * row], q1 = q[3 * row + 1], q2 = q[3 * row + 2];
      q[3 * row] = y0 + tb * q0;
      q[3 * row + 1] = y1 + tb * q1;
      q[3 * row + 2] = y2 + tb * q2;
      dl += (double)z[3 * row] * (double)y0 + (double)z[3 * row + 1] * (double)y1 +
            (double)z[3 * row + 2] * (double)y2;
    }
  }
  const double s = block_sum<SELL_BS>(dl, red);
  if (threadIdx.x == 0) A.part1[blockIdx.x] = s;
}

// One thread per source block, in BSR3 order (coalesced reads); a slice's
// writes land together in L2.  Padding slots are never written: a lane
// stops at its row's length and never reads them.
template <class T>
__global__ __launch_bounds__(256) void sell_pack_kernel(long nb, const int* __restrict__ brow,
                                                        const long* __restrict__ rowptr,
                                                        const long* __restrict__ so,
                                                        const T* __restrict__ bval,
                                                        T* __restrict__ sval) {
  for (long b = (long)blockIdx.x * 256 + threadIdx.x; b < nb; b += (long)gridDim.x * 256) {
    const int r = brow[b];
    const long j = b - rowptr[r];
    T* V = sval + (so[r >> 6] + 64 * j) * 9 + (r & 63);
    T v[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) v[q] = bval[9 * b + q];
#pragma unroll
    for (int q = 0; q < 9; ++q) V[64 * q] = v[q];
  }
}

void launch_sell_pack(hipStream_t s, bool fp32, long nb, const int* brow, const long* rowptr,
                      const long* so, const void* bval, void* sval) {
  const int g = (int)std::min<long>((nb + 255) / 256, 4096);
  if (g <= 0) return;
  if (fp32)
    hipLaunchKernelGGL(sell_pack_kernel<float>, dim3(g), dim3(256), 0, s, nb, brow, rowptr, so,
                       reinterpret_cast<const float*>(bval), reinterpret_cast<float*>(sval));
  else
    hipLaunchKernelGGL(sell_pack_kernel<double>, dim3(g), dim3(256), 0, s, nb, brow, rowptr, so,
                       reinterpret_cast<const double*>(bval), reinterpret_cast<double*>(sval));
}

template <class T, int NPT>
struct Cg1Batch {
  N3<T> Z[NPT], P[NPT], W[NPT], Q[NPT], X[NPT], R[NPT];
  PreM<T> M[NPT];
};

template <class T, bool QF, int NPT>
__device__ __forceinline__ void cg1_load(const PcgArgs& A, long i0, long hi,
                                         Cg1Batch<T, NPT>& b) {
#pragma unroll
  for (int j = 0; j < NPT; ++j) {
    const long i = i0 + (long)j * K2_BS;
    if (i < hi) {
      b.Z[j] = ld3(reinterpret_cast<const T*>(A.z), i);
      b.P[j] = ld3(reinterpret_cast<const T*>(A.p0), i);
      if (!QF) b.W[j] = ld3(reinterpret_cast<const T*>(A.p1), i);
      b.Q[j] = ld3(reinterpret_cast<const T*>(A.q), i);
      b.X[j] = ld3(reinterpret_cast<const T*>(A.x), i);
      b.R[j] = ld3(reinterpret_cast<const T*>(A.r), i);
      ld_precond<T>(A.precond, reinterpret_cast<const T*>(A.dinv),
                    reinterpret_cast<const T*>(A.binv), i, b.M[j]);
    }
  }
}

// QF: the SpMV launch has formed q = w + beta q already and left beta and
// r.z in cgs[4], cgs[5]; w is not read and q not written here.
template <class T, bool QF, int NPT>
__global__ __launch_bounds__(K2_BS) void cg1_b_kernel(PcgArgs A, int k) {
  k = chunk_k(A, k);
  __shared__ double red[3 * K2_BS / 64 + 4];
  const RowRange rg = block_rows(A.n);
  // the first batch, the flag and the partial sums are loaded together
  // (all in bounds whether or not the solve has ended): the latencies overlap
  Cg1Batch<T, NPT> b;
  long i0 = rg.lo + threadIdx.x;
  cg1_load<T, QF, NPT>(A, i0, rg.hi, b);
  const int done = A.sc->done;
  const double cdot = (!QF && A.coarse) ? *A.cg.cdot : 0.0;
  double gam, del, unused;
  all_sum3<K2_BS>(rr_slot(A, k - 1) + 1, QF ? 0 : rr_count(A, k - 1), 2, A.part1, A.g1, 1,
                  A.part1, 0, 1, red, gam, del, unused);
  if (done) return;
  double alpha, beta;
  if (QF) gam = A.cgs[5];
  else gam += cdot;
  if (k == 0) {
    beta = 0.0;
    alpha = gam / del;
  } else {
    const double* prev = A.cgs + 2 * ((k - 1) & 1);
    beta = QF ? A.cgs[4] : gam / prev[1];
    alpha = gam / (del - beta * gam / prev[0]);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    A.cgs[2 * (k & 1)] = alpha;
    A.cgs[2 * (k & 1) + 1] = gam;
  }
  const T ta = (T)alpha, tb = (T)beta;
  T* p = reinterpret_cast<T*>(A.p0);
  T* q = reinterpret_cast<T*>(A.q);
  T* x = reinterpret_cast<T*>(A.x);
  T* r = reinterpret_cast<T*>(A.r);
  T* z = reinterpret_cast<T*>(A.z);
  double rr = 0.0, rz = 0.0;
  for (;;) {
#pragma unroll
    for (int j = 0; j < NPT; ++j) {
      const long i = i0 + (long)j * K2_BS;
      if (i < rg.hi) {
        N3<T> pi, qi, xi, ri;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          pi.c[c] = b.Z[j].c[c] + tb * b.P[j].c[c];
          qi.c[c] = QF ? b.Q[j].c[c] : b.W[j].c[c] + tb * b.Q[j].c[c];
          xi.c[c] = b.X[j].c[c] + ta * pi.c[c];
          ri.c[c] = b.R[j].c[c] - ta * qi.c[c];
        }
        const N3<T> zi = mul_precond<T>(A.precond, b.M[j], ri);
        st3(p, i, pi);
        if (!QF) st3(q, i, qi);
        st3(x, i, xi);
        st3(r, i, ri);
        st3(z, i, zi);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          rr += (double)ri.c[c] * (double)ri.c[c];
          rz += (double)ri.c[c] * (double)zi.c[c];
        }
      }
    }
    i0 += (long)NPT * K2_BS;
    if (i0 >= rg.hi) break;
    cg1_load<T, QF, NPT>(A, i0, rg.hi, b);
  }
  block_sum2<K2_BS>(rr, rz, red);
  if (threadIdx.x == 0) {
    double* o = A.part2 + (long)(k & 1) * 2 * PCG_MAXP;
    o[2 * blockIdx.x] = rr;
    o[2 * blockIdx.x + 1] = rz;
  }
}

// ---- K2: x += alpha p; r -= alpha q; z = M^-1 r; r.r, r.z ---------------------
template <class T>
struct K2Batch {
  N3<T> P[K2_NPT], Q[K2_NPT], R[K2_NPT], X[K2_NPT];
  PreM<T> M[K2_NPT];
};

template <class T>
__device__ __forceinline__ void k2_load(const PcgArgs& A, const T* pk, long i0, long hi,
                                        K2Batch<T>& b) {
#pragma unroll
  for (int j = 0; j < K2_NPT; ++j) {
    const long i = i0 + (long)j * K2_BS;
    if (i < hi) {
      b.P[j] = ld3(pk, i);
      b.Q[j] = ld3(reinterpret_cast<const T*>(A.q), i);
      b.R[j] = ld3(reinterpret_cast<const T*>(A.r), i);
      b.X[j] = ld3(reinterpret_cast<const T*>(A.x), i);
      ld_precond<T>(A.precond, reinterpret_cast<const T*>(A.dinv),
                    reinterpret_cast<const T*>(A.binv), i, b.M[j]);
    }
  }
}

template <class T>
__global__ __launch_bounds__(K2_BS) void pcg_k2_kernel(PcgArgs A, int k) {
  k = chunk_k(A, k);
  __shared__ double red[3 * K2_BS / 64 + 4];
  PcgScalars* sc = A.sc;
  if (sc->done) return;
  const T* pk = reinterpret_cast<const T*>((k & 1) ? A.p1 : A.p0);
  T* x = reinterpret_cast<T*>(A.x);
  T* r = reinterpret_cast<T*>(A.r);
  T* z = reinterpret_cast<T*>(A.z);
  const RowRange rg = block_rows(A.n);
  // the first batch is loaded before the partial sums: the two latencies overlap
  K2Batch<T> b;
  long i0 = rg.lo + threadIdx.x;
  k2_load<T>(A, pk, i0, rg.hi, b);
  double pq, rz, unused;
  all_sum3<K2_BS>(A.part1, A.g1, 1, rr_slot(A, k - 1) + 1, rr_count(A, k - 1), 2, A.part1, 0, 1,
                  red, pq, rz, unused);
  const T ta = (T)(rz / pq);
  double rr = 0.0, rzn = 0.0;
  for (;;) {
#pragma unroll
    for (int j = 0; j < K2_NPT; ++j) {
      const long i = i0 + (long)j * K2_BS;
      if (i < rg.hi) {
        N3<T> xi, ri;
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          xi.c[p] = b.X[j].c[p] + ta * b.P[j].c[p];
          ri.c[p] = b.R[j].c[p] - ta * b.Q[j].c[p];
        }
        const N3<T> zi = mul_precond<T>(A.precond, b.M[j], ri);
        st3(x, i, xi);
        st3(r, i, ri);
        st3(z, i, zi);
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          rr += (double)ri.c[p] * (double)ri.c[p];
          rzn += (double)ri.c[p] * (double)zi.c[p];
        }
      }
    }
    i0 += (long)K2_NPT * K2_BS;
    if (i0 >= rg.hi) break;
    k2_load<T>(A, pk, i0, rg.hi, b);
  }
  block_sum2<K2_BS>(rr, rzn, red);
  if (threadIdx.x == 0) {
    double* o = A.part2 + (long)(k & 1) * 2 * PCG_MAXP;
    o[2 * blockIdx.x] = rr;
    o[2 * blockIdx.x + 1] = rzn;
  }
}

// ---- status after iterations [0, k_next) ---------------------------------------
__global__ __launch_bounds__(256) void pcg_status_kernel(PcgArgs A, int k_next, int maxit) {
  if (A.kctl) {
    k_next += A.kctl[0];
    maxit = A.kctl[1];
  }
  __shared__ double red[2 * 256 / 64 + 2];
  PcgScalars* sc = A.sc;
  if (sc->done) return;
  const int k = k_next;
  const double rr = all_sum<256>(rr_slot(A, k - 1), rr_count(A, k - 1), 2, red);
  double thr;
  if (k == 0) {
    const double bb = all_sum<256>(A.partb, A.g1, 1, red);
    thr = A.tol * A.tol * bb;
    if (thr < 2.2250738585072014e-308) thr = 2.2250738585072014e-308;
    if (threadIdx.x == 0) {
      sc->bb = bb;
      sc->thr = thr;
    }
  } else {
    thr = sc->thr;
  }
  if (threadIdx.x == 0) {
    sc->rr = rr;
    if (rr < thr) {
      sc->done = 1;
      sc->iters = k > 0 ? k - 1 : 0;
    } else if (k >= maxit) {
      sc->done = 1;
      sc->iters = maxit;
    }
  }
}

template <int L>
static void launch_init_l(hipStream_t s, bool fp32, const PcgArgs& a, const void* b, int xz) {
  if (fp32)
    hipLaunchKernelGGL((pcg_init_kernel<float, L>), dim3(a.g1), dim3(K1_BS), 0, s, a,
                       reinterpret_cast<const float*>(b), xz);
  else
    hipLaunchKernelGGL((pcg_init_kernel<double, L>), dim3(a.g1), dim3(K1_BS), 0, s, a,
                       reinterpret_cast<const double*>(b), xz);
}

void launch_pcg_init(hipStream_t s, bool fp32, const PcgArgs& a, const void* b, int x_zero) {
  switch (a.lpr) {
    case 4: launch_init_l<4>(s, fp32, a, b, x_zero); break;
    case 16: launch_init_l<16>(s, fp32, a, b, x_zero); break;
    default: launch_init_l<8>(s, fp32, a, b, x_zero); break;
  }
}

// Launch with optional start / stop events taken from the dispatch itself
// (hipExtLaunchKernelGGL), so a sampled launch is timed without the gap an
// event-record command would add.
template <class K>
static void launch_timed(K kernel, dim3 grid, dim3 block, hipStream_t s, hipEvent_t e0,
                         hipEvent_t e1, const PcgArgs& a, int k) {
  if (e0)
    hipExtLaunchKernelGGL(kernel, grid, block, 0, s, e0, e1, 0, a, k);
  else
    hipLaunchKernelGGL(kernel, grid, block, 0, s, a, k);
}

template <int L>
static void launch_k1_l(hipStream_t s, bool fp32, const PcgArgs& a, int k, hipEvent_t e0,
                        hipEvent_t e1) {
  if (a.cg1 && a.sell) {
    if (a.sell_c16) {
      if (fp32)
        launch_timed(cg1_sell_kernel<float, true>, dim3(a.g1), dim3(SELL_BS), s, e0, e1, a, k);
      else
        launch_timed(cg1_sell_kernel<double, true>, dim3(a.g1), dim3(SELL_BS), s, e0, e1, a, k);
    } else {
      if (fp32)
        launch_timed(cg1_sell_kernel<float, false>, dim3(a.g1), dim3(SELL_BS), s, e0, e1, a, k);
      else
        launch_timed(cg1_sell_kernel<double, false>, dim3(a.g1), dim3(SELL_BS), s, e0, e1, a,
                     k);
    }
    return;
  }
  if (a.cg1) {
    if (fp32)
      launch_timed(cg1_a_kernel<float, L>, dim3(a.g1), dim3(K1_BS), s, e0, e1, a, k);
    else
      launch_timed(cg1_a_kernel<double, L>, dim3(a.g1), dim3(K1_BS), s, e0, e1, a, k);
    return;
  }
  if (a.split) {
    if (a.mf) {  // matrix-free product on this iteration's p (direct launches only)
      if (e0) PGN_HIP(hipEventRecord(e0, s));
      launch_matfree_product(s, fp32, a.mfa, (k & 1) ? a.p1 : a.p0, a.q, a.part1, a.g1);
      if (e1) PGN_HIP(hipEventRecord(e1, s));
      return;
    }
    if (fp32)
      launch_timed(pcg_k1b_kernel<float, L>, dim3(a.g1), dim3(K1_BS), s, e0, e1, a, k);
    else
      launch_timed(pcg_k1b_kernel<double, L>, dim3(a.g1), dim3(K1_BS), s, e0, e1, a, k);
    return;
  }
  if (fp32)
    launch_timed(pcg_k1_kernel<float, L>, dim3(a.g1), dim3(K1_BS), s, e0, e1, a, k);
  else
    launch_timed(pcg_k1_kernel<double, L>, dim3(a.g1), dim3(K1_BS), s, e0, e1, a, k);
}

void launch_pcg_k1a(hipStream_t s, bool fp32, const PcgArgs& a, int k) {
  if (!a.split || a.cg1) return;
  if (fp32)
    hipLaunchKernelGGL(pcg_k1a_kernel<float>, dim3(a.g2), dim3(K2_BS), 0, s, a, k);
  else
    hipLaunchKernelGGL(pcg_k1a_kernel<double>, dim3(a.g2), dim3(K2_BS), 0, s, a, k);
}

void launch_pcg_k1(hipStream_t s, bool fp32, const PcgArgs& a, int k, hipEvent_t e0,
                   hipEvent_t e1) {
  switch (a.lpr) {
    case 4: launch_k1_l<4>(s, fp32, a, k, e0, e1); break;
    case 16: launch_k1_l<16>(s, fp32, a, k, e0, e1); break;
    default: launch_k1_l<8>(s, fp32, a, k, e0, e1); break;
  }
}

void launch_pcg_k2(hipStream_t s, bool fp32, const PcgArgs& a, int k, hipEvent_t e0,
                   hipEvent_t e1) {
  if (a.cg1) {
    if (a.sell) {
      if (fp32)
        launch_timed(cg1_b_kernel<float, true, K2_NPT>, dim3(a.g2), dim3(K2_BS), s, e0, e1, a,
                     k);
      else
        launch_timed(cg1_b_kernel<double, true, K2_NPT>, dim3(a.g2), dim3(K2_BS), s, e0, e1, a,
                     k);
    } else {
      if (fp32)
        launch_timed(cg1_b_kernel<float, false, K2_NPT>, dim3(a.g2), dim3(K2_BS), s, e0, e1, a,
                     k);
      else
        launch_timed(cg1_b_kernel<double, false, K2_NPT>, dim3(a.g2), dim3(K2_BS), s, e0, e1,
                     a, k);
    }
    return;
  }
  if (fp32)
    launch_timed(pcg_k2_kernel<float>, dim3(a.g2), dim3(K2_BS), s, e0, e1, a, k);
  else
    launch_timed(pcg_k2_kernel<double>, dim3(a.g2), dim3(K2_BS), s, e0, e1, a, k);
}

void launch_pcg_status(hipStream_t s, const PcgArgs& a, int k_next, int maxit) {
  hipLaunchKernelGGL(pcg_status_kernel, dim3(1), dim3(256), 0, s, a, k_next, maxit);
}

// ---- FP64 residual with a precision-T matrix --------------------------------
template <class TM>
__global__ __launch_bounds__(PCG_BS) void residual64_kernel(
    int n, const long* __restrict__ rowptr, const int* __restrict__ bcol,
    const TM* __restrict__ bval, const double* __restrict__ x,
    const double* __restrict__ b, double* __restrict__ r, double* part,
    double* out, unsigned* ticket, long row0) {
  __shared__ double red[PCG_BS / 64];
  __shared__ bool last;
  const int lane = threadIdx.x & (LPR - 1);
  double rr = 0.0;
  const long row1 = row0 + n;
  for (long row = row0 + (blockIdx.x * PCG_BS + threadIdx.x) / LPR; row < row1;
       row += gridDim.x * ROWS_PER_BLOCK) {
    double y0 = 0, y1 = 0, y2 = 0;
    const long e1 = rowptr[row + 1];
    for (long s = rowptr[row] + lane; s < e1; s += LPR) {
      const long j = bcol[s];
      const TM* B = bval + s * 9;
      const double xj0 = x[3 * j], xj1 = x[3 * j + 1], xj2 = x[3 * j + 2];
      y0 += (double)B[0] * xj0 + (double)B[1] * xj1 + (double)B[2] * xj2;
      y1 += (double)B[3] * xj0 + (double)B[4] * xj1 + (double)B[5] * xj2;
      y2 += (double)B[6] * xj0 + (double)B[7] * xj1 + (double)B[8] * xj2;
    }
    y0 = group_sum8(y0);
    y1 = group_sum8(y1);
    y2 = group_sum8(y2);
    if (lane == 0) {
      const long i = row;
      const double r0 = b[3 * i] - y0, r1 = b[3 * i + 1] - y1, r2 = b[3 * i + 2] - y2;
      r[3 * i] = r0;
      r[3 * i + 1] = r1;
      r[3 * i + 2] = r2;
      rr += r0 * r0 + r1 * r1 + r2 * r2;
    }
  }
  const double s = block_sum<PCG_BS>(rr, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if (last_block(ticket, &last)) {
    const double t = block_sum_array<PCG_BS>(part, gridDim.x, 1, red);
    if (threadIdx.x == 0) out[0] = t;
  }
}

// The same residual over the SELL-64 copy of the matrix (one row per lane,
// the row's blocks in column order; grid and slice walk of cg1_sell_kernel).
template <class TM, bool C16>
__global__ __launch_bounds__(SELL_BS) void residual64_sell_kernel(
    PcgArgs A, const double* __restrict__ x, const double* __restrict__ b,
    double* __restrict__ r, double* part, double* out, unsigned* ticket) {
  __shared__ double red[SELL_BS / 64];
  __shared__ bool last;
  constexpr int WAVES = SELL_BS / 64;
  const TM* __restrict__ sval = reinterpret_cast<const TM*>(A.sell_val);
  const int* __restrict__ scol = A.sell_col;
  const short* __restrict__ s16 = A.sell_c16;
  const int G = gridDim.x, bk = blockIdx.x;
  const int lb = (G % 8 == 0) ? (bk % 8) * (G / 8) + bk / 8 : bk;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double rr = 0.0;
  for (int t = wave; t < A.sell_spw; t += WAVES) {
    const long slice = (long)lb * A.sell_spw + t;
    if (slice >= A.sell_ns) break;
    const long row = slice * 64 + lane;
    if (row >= A.n) continue;
    const int len = A.sell_len[row];
    const long base = A.sell_so[slice];
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
    for (int j = 0; j < len; ++j) {
      const long c = C16 ? row + s16[base + 64L * j + lane] : scol[base + 64L * j + lane];
      const TM* V = sval + (base + 64L * j) * 9 + lane;
      const double a0 = x[3 * c], a1 = x[3 * c + 1], a2 = x[3 * c + 2];
      y0 += (double)V[0] * a0 + (double)V[64] * a1 + (double)V[128] * a2;
      y1 += (double)V[192] * a0 + (double)V[256] * a1 + (double)V[320] * a2;
      y2 += (double)V[384] * a0 + (double)V[448] * a1 + (double)V[512] * a2;
    }
    const double r0 = b[3 * row] - y0, r1 = b[3 * row + 1] - y1, r2 = b[3 * row + 2] - y2;
    r[3 * row] = r0;
    r[3 * row + 1] = r1;
    r[3 * row + 2] = r2;
    rr += r0 * r0 + r1 * r1 + r2 * r2;
  }
  const double sum = block_sum<SELL_BS>(rr, red);
  if (threadIdx.x == 0) part[blockIdx.x] = sum;
  if (last_block(ticket, &last)) {
    const double t = block_sum_array<SELL_BS>(part, gridDim.x, 1, red);
    if (threadIdx.x == 0) out[0] = t;
  }
}

void launch_residual64_sell(hipStream_t s, bool fp32_mat, const PcgArgs& a, const double* x,
                            const double* b, double* r, double* partials, double* out) {
  unsigned* ticket = reinterpret_cast<unsigned*>(partials + PCG_MAXP);
  const dim3 g(a.g1), bs(SELL_BS);
  if (fp32_mat) {
    if (a.sell_c16)
      hipLaunchKernelGGL((residual64_sell_kernel<float, true>), g, bs, 0, s, a, x, b, r,
                         partials, out, ticket);
    else
      hipLaunchKernelGGL((residual64_sell_kernel<float, false>), g, bs, 0, s, a, x, b, r,
                         partials, out, ticket);
  } else {
    if (a.sell_c16)
      hipLaunchKernelGGL((residual64_sell_kernel<double, true>), g, bs, 0, s, a, x, b, r,
                         partials, out, ticket);
    else
      hipLaunchKernelGGL((residual64_sell_kernel<double, false>), g, bs, 0, s, a, x, b, r,
                         partials, out, ticket);
  }
}

// partials: PCG_MAXP doubles followed by an unsigned ticket
void launch_residual64(hipStream_t s, bool fp32_mat, int n, const long* rowptr,
                       const int* bcol, const void* bval, const double* x,
                       const double* b, double* r, double* partials,
                       double* out, long row0) {
  if (n <= 0) return;
  const int g = grid_rows(n);
  unsigned* ticket = reinterpret_cast<unsigned*>(partials + PCG_MAXP);
  if (fp32_mat)
    hipLaunchKernelGGL(residual64_kernel<float>, dim3(g), dim3(PCG_BS), 0, s, n, rowptr,
                       bcol, reinterpret_cast<const float*>(bval), x, b, r, partials,
                       out, ticket, row0);
  else
    hipLaunchKernelGGL(residual64_kernel<double>, dim3(g), dim3(PCG_BS), 0, s, n, rowptr,
                       bcol, reinterpret_cast<const double*>(bval), x, b, r, partials,
                       out, ticket, row0);
}

// ---- plain SpMV (tuning, timing) --------------------------------------------
template <class T, int L>
__global__ __launch_bounds__(K1_BS) void spmv_kernel(int n, const long* __restrict__ rowptr,
                                                    const int* __restrict__ bcol,
                                                    const T* __restrict__ bval,
                                                    const T* __restrict__ x,
                                                    T* __restrict__ y) {
  const int lane = threadIdx.x & (L - 1);
  const RowRange rg = block_rows(n);
  for (long row = rg.lo + threadIdx.x / L; row < rg.hi; row += K1_BS / L) {
    T y0 = 0, y1 = 0, y2 = 0;
    const long e1 = rowptr[row + 1];
    for (long s = rowptr[row] + lane; s < e1; s += L) {
      const long j = bcol[s];
      const T* B = bval + s * 9;
      const T xj0 = x[3 * j], xj1 = x[3 * j + 1], xj2 = x[3 * j + 2];
      y0 += B[0] * xj0 + B[1] * xj1 + B[2] * xj2;
      y1 += B[3] * xj0 + B[4] * xj1 + B[5] * xj2;
      y2 += B[6] * xj0 + B[7] * xj1 + B[8] * xj2;
    }
    y0 = group_sum<L>(y0);
    y1 = group_sum<L>(y1);
    y2 = group_sum<L>(y2);
    if (lane == L - 1) {
      y[3L * row] = y0;
      y[3L * row + 1] = y1;
      y[3L * row + 2] = y2;
    }
  }
}

template <int L>
static void launch_spmv_l(hipStream_t s, bool fp32, int n, const long* rowptr, const int* bcol,
                          const void* bval, const void* x, void* y) {
  const int g = pcg_grid1(n, L);
  if (fp32)
    hipLaunchKernelGGL((spmv_kernel<float, L>), dim3(g), dim3(K1_BS), 0, s, n, rowptr, bcol,
                       reinterpret_cast<const float*>(bval),
                       reinterpret_cast<const float*>(x), reinterpret_cast<float*>(y));
  else
    hipLaunchKernelGGL((spmv_kernel<double, L>), dim3(g), dim3(K1_BS), 0, s, n, rowptr, bcol,
                       reinterpret_cast<const double*>(bval),
                       reinterpret_cast<const double*>(x), reinterpret_cast<double*>(y));
}

void launch_spmv(hipStream_t s, bool fp32, int n, const long* rowptr, const int* bcol,
                 const void* bval, const void* x, void* y, int lpr) {
  switch (lpr) {
    case 4: launch_spmv_l<4>(s, fp32, n, rowptr, bcol, bval, x, y); break;
    case 16: launch_spmv_l<16>(s, fp32, n, rowptr, bcol, bval, x, y); break;
    default: launch_spmv_l<8>(s, fp32, n, rowptr, bcol, bval, x, y); break;
  }
}

// ---- vector helpers -------------------------------------------------------------
__global__ void cast_kernel(long n, const double* __restrict__ a, float* __restrict__ b) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    b[i] = (float)a[i];
}
__global__ void copy64_kernel(long n, const double* __restrict__ a, double* __restrict__ b) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    b[i] = a[i];
}
__global__ void axpy_f32_kernel(long n, const float* __restrict__ d, double* __restrict__ x) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    x[i] += (double)d[i];
}
__global__ void lincomb64_kernel(long n, double a, const double* x, double b,
                                 const double* y, double* out) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    out[i] = a * x[i] + b * y[i];
}
template <class T>
__global__ void widen_kernel(long n, const T* __restrict__ a, double* __restrict__ b) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    b[i] = (double)a[i];
}

__global__ void copy_f32_kernel(long n4, const float4* __restrict__ a, float4* __restrict__ b) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n4; i += (long)gridDim.x * 256)
    b[i] = a[i];
}

static inline int vgrid(long n) {
  long g = (n + 255) / 256;
  return (int)(g < 1 ? 1 : (g > 2048 ? 2048 : g));
}

void launch_copy_f32(hipStream_t s, long n, const float* a, float* b) {
  hipLaunchKernelGGL(copy_f32_kernel, dim3(8192), dim3(256), 0, s, n / 4,
                     reinterpret_cast<const float4*>(a), reinterpret_cast<float4*>(b));
}

void launch_cast_f64_to_f32(hipStream_t s, long n, const double* a, float* b) {
  hipLaunchKernelGGL(cast_kernel, dim3(vgrid(n)), dim3(256), 0, s, n, a, b);
}
void launch_axpy_f32_to_f64(hipStream_t s, long n, const float* d, double* x) {
  hipLaunchKernelGGL(axpy_f32_kernel, dim3(vgrid(n)), dim3(256), 0, s, n, d, x);
}
void launch_lincomb64(hipStream_t s, long n, double a, const double* x, double b,
                      const double* y, double* out) {
  hipLaunchKernelGGL(lincomb64_kernel, dim3(vgrid(n)), dim3(256), 0, s, n, a, x, b, y, out);
}
void launch_widen(hipStream_t s, bool fp32, long n, const void* a, double* b) {
  if (fp32)
    hipLaunchKernelGGL(widen_kernel<float>, dim3(vgrid(n)), dim3(256), 0, s, n,
                       reinterpret_cast<const float*>(a), b);
  else
    hipLaunchKernelGGL(copy64_kernel, dim3(vgrid(n)), dim3(256), 0, s, n,
                       reinterpret_cast<const double*>(a), b);
}
void launch_narrow_or_copy(hipStream_t s, bool fp32, long n, const double* a, void* b) {
  if (fp32)
    hipLaunchKernelGGL(cast_kernel, dim3(vgrid(n)), dim3(256), 0, s, n, a,
                       reinterpret_cast<float*>(b));
  else
    hipLaunchKernelGGL(copy64_kernel, dim3(vgrid(n)), dim3(256), 0, s, n, a,
                       reinterpret_cast<double*>(b));
}

constexpr int DOT_GRID = 512;
__global__ __launch_bounds__(256) void dot64_kernel(long n, const double* __restrict__ a,
                                                   const double* __restrict__ b,
                                                   double* part, double* out,
                                                   unsigned* ticket) {
  __shared__ double red[4];
  __shared__ bool last;
  double acc = 0.0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    acc += a[i] * b[i];
  const double s = block_sum<256>(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if (last_block(ticket, &last)) {
    const double t = block_sum_array<256>(part, gridDim.x, 1, red);
    if (threadIdx.x == 0) out[0] = t;
  }
}

// partials: PCG_MAXP doubles followed by an unsigned ticket
void launch_dot64(hipStream_t s, long n, const double* a, const double* b,
                  double* partials, double* out) {
  unsigned* ticket = reinterpret_cast<unsigned*>(partials + PCG_MAXP);
  hipLaunchKernelGGL(dot64_kernel, dim3(DOT_GRID), dim3(256), 0, s, n, a, b, partials,
                     out, ticket);
}

}  // namespace pgn

// ---- segmented PCG -------------------------------------------------------------
// Same iteration as above, one independent PCG per segment (instance).  A
// workgroup serves one segment: it walks that segment's rows with the
// segment's own stride (G workgroups), writes its partials at its slot, and
// re-sums only its own segment's G partials.
namespace pgn {

constexpr int SEG_K2_BS = 256;

int seg_rows_per_wg(int lpr) { return K1_BS / lpr; }

__device__ __forceinline__ const double* seg_rr_slot(const SegPcgArgs& A, int k) {
  return A.part2 + (long)(k & 1) * 2 * A.W;
}
// r.(P uc) of iteration k's coarse correction (two-level only)
__device__ __forceinline__ double seg_cdot(const SegPcgArgs& A, int k, int sg) {
  return A.coarse ? A.cg.cdot[(long)(k & 1) * A.K + sg] : 0.0;
}

template <class T, int L>
__global__ __launch_bounds__(K1_BS) void seg_init_kernel(SegPcgArgs A, const T* __restrict__ b,
                                                         int x_zero) {
  __shared__ double red[3 * K1_BS / 64];
  const int w = blockIdx.x, sg = A.wg_seg[w], j = A.wg_loc[w], G = A.seg_G[sg];
  const long r0 = A.seg_row0[sg], r1 = A.seg_row1[sg];
  const T* __restrict__ bval = reinterpret_cast<const T*>(A.bval);
  T* __restrict__ x = reinterpret_cast<T*>(A.x);
  T* __restrict__ r = reinterpret_cast<T*>(A.r);
  T* __restrict__ z = reinterpret_cast<T*>(A.z);
  T* __restrict__ p0 = reinterpret_cast<T*>(A.p0);
  T* __restrict__ p1 = reinterpret_cast<T*>(A.p1);
  const int lane = threadIdx.x & (L - 1);
  constexpr int RPW = K1_BS / L;
  double bb = 0.0, rr = 0.0, rz = 0.0;
  for (long row = r0 + (long)j * RPW + threadIdx.x / L; row < r1; row += (long)G * RPW) {
    T y0 = 0, y1 = 0, y2 = 0;
    if (!x_zero) {
      const long e1 = A.rowptr[row + 1];
      for (long s = A.rowptr[row] + lane; s < e1; s += L) {
        const long c = A.bcol[s];
        const T* B = bval + s * 9;
        const T xj0 = x[3 * c], xj1 = x[3 * c + 1], xj2 = x[3 * c + 2];
        y0 += B[0] * xj0 + B[1] * xj1 + B[2] * xj2;
        y1 += B[3] * xj0 + B[4] * xj1 + B[5] * xj2;
        y2 += B[6] * xj0 + B[7] * xj1 + B[8] * xj2;
      }
      y0 = group_sum<L>(y0);
      y1 = group_sum<L>(y1);
      y2 = group_sum<L>(y2);
    }
    if (lane == L - 1) {
      const long i = row;
      T ri[3] = {b[3 * i] - y0, b[3 * i + 1] - y1, b[3 * i + 2] - y2};
      T zi[3];
      apply_precond<T>(A.precond, reinterpret_cast<const T*>(A.dinv),
                       reinterpret_cast<const T*>(A.binv), i, ri, zi);
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        r[3 * i + p] = ri[p];
        z[3 * i + p] = zi[p];
        p0[3 * i + p] = T(0);
        p1[3 * i + p] = T(0);
        if (x_zero) x[3 * i + p] = T(0);
        bb += (double)b[3 * i + p] * (double)b[3 * i + p];
        rr += (double)ri[p] * (double)ri[p];
        rz += (double)ri[p] * (double)zi[p];
      }
    }
  }
  const double sbb = block_sum<K1_BS>(bb, red);
  block_sum2<K1_BS>(rr, rz, red);
  if (threadIdx.x == 0) {
    A.partb[w] = sbb;
    double* o = A.part2 + 2L * A.W;  // parity slot of iteration -1
    o[2 * w] = rr;
    o[2 * w + 1] = rz;
    if (j == 0) {
      A.sc[sg].done = 0;
      A.sc[sg].iters = 0;
    }
  }
}

// z_i + F_i (t_a + w_a x rel_i): the two-level z of node i from the stored
// block-Jacobi part and the last coarse solution uc (the prolongation is
// applied where z is read instead of being stored)
// ucl: the instance's uc staged in LDS, a0 its first aggregate
template <class T>
__device__ __forceinline__ void seg_coarse_z(const SegCoarseArgs& c, const double* ucl, int a0,
                                             long i, T& z0, T& z1, T& z2) {
  const T* __restrict__ rel = reinterpret_cast<const T*>(c.rel);
  const int a = c.agg_of[i] - a0;
  const unsigned fm = c.fixmask[i];
  const double x = rel[3 * i], y = rel[3 * i + 1], zz = rel[3 * i + 2];
  const double* u = ucl + 6 * a;
  const double d0 = u[0] + u[4] * zz - u[5] * y;
  const double d1 = u[1] + u[5] * x - u[3] * zz;
  const double d2 = u[2] + u[3] * y - u[4] * x;
  if (!(fm & 1u)) z0 = (T)((double)z0 + d0);
  if (!(fm & 2u)) z1 = (T)((double)z1 + d1);
  if (!(fm & 4u)) z2 = (T)((double)z2 + d2);
}

template <class T, int L, bool C16, bool CG>
__global__ __launch_bounds__(K1_BS) void seg_k1_kernel(SegPcgArgs A, int k) {
  k = chunk_k(A, k);
  __shared__ double red[3 * K1_BS / 64 + 3];
  const int w = blockIdx.x, sg = A.wg_seg[w], j = A.wg_loc[w], G = A.seg_G[sg];
  PcgScalars* sc = A.sc + sg;
  if (sc->done) return;
  const int g0 = A.seg_wg0[sg];
  const double* prev = seg_rr_slot(A, k - 1) + 2L * g0;
  // the instance's coarse solution in LDS (two-level), loaded while the
  // partial sums are read
  __shared__ double ucl[CG ? 6 * SEG_COARSE_MAXAGG : 1];
  int ca0 = 0;
  if (CG) {
    ca0 = A.cg.inst_aoff[sg];
    const int nc = 6 * (A.cg.inst_aoff[sg + 1] - ca0);
    for (int e = threadIdx.x; e < nc; e += K1_BS) ucl[e] = A.cg.uc[6L * ca0 + e];
  }
  // r.r, r.z of iteration k - 1 and b.b (k = 0) or r.z of k - 2 in one pass
  double rr, rz, third;
  if (k == 0)
    all_sum3<K1_BS>(prev, G, 2, prev + 1, G, 2, A.partb + g0, G, 1, red, rr, rz, third);
  else
    all_sum3<K1_BS>(prev, G, 2, prev + 1, G, 2, seg_rr_slot(A, k - 2) + 2L * g0 + 1, G, 2, red,
                    rr, rz, third);
  rz += seg_cdot(A, k - 1, sg);
  double thr, beta;
  if (k == 0) {
    const double bb = third;
    thr = A.tol[sg] * A.tol[sg] * bb;
    if (thr < 2.2250738585072014e-308) thr = 2.2250738585072014e-308;
    if (j == 0 && threadIdx.x == 0) {
      sc->thr = thr;
      sc->bb = bb;
      sc->rr = rr;
    }
    if (bb == 0.0 || rr < thr) {
      if (j == 0 && threadIdx.x == 0) {
        sc->done = 1;
        sc->iters = 0;
      }
      return;
    }
    beta = 0.0;
  } else {
    thr = sc->thr;
    if (rr < thr) {
      if (j == 0 && threadIdx.x == 0) {
        sc->done = 1;
        sc->iters = k - 1;
        sc->rr = rr;
      }
      return;
    }
    const double rzm = third + seg_cdot(A, k - 2, sg);
    beta = rz / rzm;
  }
  const T tb = (T)beta;
  const T* __restrict__ bval = reinterpret_cast<const T*>(A.bval);
  const T* __restrict__ z = reinterpret_cast<const T*>(A.z);
  T* __restrict__ pnew = reinterpret_cast<T*>((k & 1) ? A.p1 : A.p0);
  const T* __restrict__ pold = reinterpret_cast<const T*>((k & 1) ? A.p0 : A.p1);
  T* __restrict__ q = reinterpret_cast<T*>(A.q);
  const int lane = threadIdx.x & (L - 1);
  constexpr int RPW = K1_BS / L;
  const long r0 = A.seg_row0[sg], r1 = A.seg_row1[sg];
  double pq = 0.0;
  for (long row = r0 + (long)j * RPW + threadIdx.x / L; row < r1; row += (long)G * RPW) {
    T y0 = 0, y1 = 0, y2 = 0;
    const long e1 = A.rowptr[row + 1];
    for (long s = A.rowptr[row] + lane; s < e1; s += L) {
      const long c = C16 ? row + A.bcol16[s] : A.bcol[s];
      const T* B = bval + s * 9;
      T zj0 = z[3 * c], zj1 = z[3 * c + 1], zj2 = z[3 * c + 2];
      if (CG) seg_coarse_z<T>(A.cg, ucl, ca0, c, zj0, zj1, zj2);
      const T pj0 = zj0 + tb * pold[3 * c];
      const T pj1 = zj1 + tb * pold[3 * c + 1];
      const T pj2 = zj2 + tb * pold[3 * c + 2];
      y0 += B[0] * pj0 + B[1] * pj1 + B[2] * pj2;
      y1 += B[3] * pj0 + B[4] * pj1 + B[5] * pj2;
      y2 += B[6] * pj0 + B[7] * pj1 + B[8] * pj2;
    }
    y0 = group_sum<L>(y0);
    y1 = group_sum<L>(y1);
    y2 = group_sum<L>(y2);
    if (lane == L - 1) {
      const long i = row;
      T zi0 = z[3 * i], zi1 = z[3 * i + 1], zi2 = z[3 * i + 2];
      if (CG) seg_coarse_z<T>(A.cg, ucl, ca0, i, zi0, zi1, zi2);
      const T pi0 = zi0 + tb * pold[3 * i];
      const T pi1 = zi1 + tb * pold[3 * i + 1];
      const T pi2 = zi2 + tb * pold[3 * i + 2];
      pnew[3 * i] = pi0;
      pnew[3 * i + 1] = pi1;
      pnew[3 * i + 2] = pi2;
      q[3 * i] = y0;
      q[3 * i + 1] = y1;
      q[3 * i + 2] = y2;
      pq += (double)pi0 * (double)y0 + (double)pi1 * (double)y1 + (double)pi2 * (double)y2;
    }
  }
  const double s = block_sum<K1_BS>(pq, red);
  if (threadIdx.x == 0) A.part1[w] = s;
}

template <class T>
__global__ __launch_bounds__(SEG_K2_BS) void seg_k2_kernel(SegPcgArgs A, int k) {
  k = chunk_k(A, k);
  __shared__ double red[3 * SEG_K2_BS / 64 + 3];
  const int w = blockIdx.x, sg = A.wg_seg[w], j = A.wg_loc[w], G = A.seg_G[sg];
  if (A.sc[sg].done) return;
  const int g0 = A.seg_wg0[sg];
  const T* pk = reinterpret_cast<const T*>((k & 1) ? A.p1 : A.p0);
  const T* q = reinterpret_cast<const T*>(A.q);
  T* x = reinterpret_cast<T*>(A.x);
  T* r = reinterpret_cast<T*>(A.r);
  T* z = reinterpret_cast<T*>(A.z);
  const T* dinv = reinterpret_cast<const T*>(A.dinv);
  const T* binv = reinterpret_cast<const T*>(A.binv);
  const long r1 = A.seg_row1[sg], stride = (long)G * SEG_K2_BS;
  long i = A.seg_row0[sg] + (long)j * SEG_K2_BS + threadIdx.x;
  // operands of the first node are loaded before the partial sums
  N3<T> P, Q, R, X;
  PreM<T> M;
  if (i < r1) {
    P = ld3(pk, i);
    Q = ld3(q, i);
    R = ld3<T>(r, i);
    X = ld3<T>(x, i);
    ld_precond<T>(A.precond, dinv, binv, i, M);
  }
  double pq, rz, none;
  all_sum3<SEG_K2_BS>(A.part1 + g0, G, 1, seg_rr_slot(A, k - 1) + 2L * g0 + 1, G, 2, nullptr, 0,
                      1, red, pq, rz, none);
  rz += seg_cdot(A, k - 1, sg);
  const T ta = (T)(rz / pq);
  double rr = 0.0, rzn = 0.0;
  while (i < r1) {
    N3<T> xi, ri;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      xi.c[p] = X.c[p] + ta * P.c[p];
      ri.c[p] = R.c[p] - ta * Q.c[p];
    }
    const N3<T> zi = mul_precond<T>(A.precond, M, ri);
    st3(x, i, xi);
    st3(r, i, ri);
    st3(z, i, zi);
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      rr += (double)ri.c[p] * (double)ri.c[p];
      rzn += (double)ri.c[p] * (double)zi.c[p];
    }
    i += stride;
    if (i < r1) {
      P = ld3(pk, i);
      Q = ld3(q, i);
      R = ld3<T>(r, i);
      X = ld3<T>(x, i);
      ld_precond<T>(A.precond, dinv, binv, i, M);
    }
  }
  block_sum2<SEG_K2_BS>(rr, rzn, red);
  if (threadIdx.x == 0) {
    double* o = A.part2 + (long)(k & 1) * 2 * A.W;
    o[2 * w] = rr;
    o[2 * w + 1] = rzn;
  }
}

// one workgroup per segment
__global__ __launch_bounds__(256) void seg_status_kernel(SegPcgArgs A, int k_next, int maxit) {
  if (A.kctl) {
    k_next += A.kctl[0];
    maxit = A.kctl[1];
  }
  __shared__ double red[2 * 256 / 64 + 2];
  const int sg = blockIdx.x;
  PcgScalars* sc = A.sc + sg;
  if (sc->done) return;
  const int k = k_next, g0 = A.seg_wg0[sg], G = A.seg_G[sg];
  const double rr = all_sum<256>(seg_rr_slot(A, k - 1) + 2L * g0, G, 2, red);
  double thr;
  if (k == 0) {
    const double bb = all_sum<256>(A.partb + g0, G, 1, red);
    thr = A.tol[sg] * A.tol[sg] * bb;
    if (thr < 2.2250738585072014e-308) thr = 2.2250738585072014e-308;
    if (threadIdx.x == 0) {
      sc->bb = bb;
      sc->thr = thr;
    }
  } else {
    thr = sc->thr;
  }
  if (threadIdx.x == 0) {
    sc->rr = rr;
    if (rr < thr) {
      sc->done = 1;
      sc->iters = k > 0 ? k - 1 : 0;
    } else if (k >= maxit) {
      sc->done = 1;
      sc->iters = maxit;
    }
  }
}

template <int L>
static void seg_init_l(hipStream_t s, bool fp32, const SegPcgArgs& a, const void* b, int xz) {
  if (fp32)
    hipLaunchKernelGGL((seg_init_kernel<float, L>), dim3(a.W), dim3(K1_BS), 0, s, a,
                       reinterpret_cast<const float*>(b), xz);
  else
    hipLaunchKernelGGL((seg_init_kernel<double, L>), dim3(a.W), dim3(K1_BS), 0, s, a,
                       reinterpret_cast<const double*>(b), xz);
}

void launch_seg_init(hipStream_t s, bool fp32, const SegPcgArgs& a, const void* b, int x_zero) {
  switch (a.lpr) {
    case 4: seg_init_l<4>(s, fp32, a, b, x_zero); break;
    case 16: seg_init_l<16>(s, fp32, a, b, x_zero); break;
    default: seg_init_l<8>(s, fp32, a, b, x_zero); break;
  }
}

template <int L, bool CG>
static void seg_k1_lc(hipStream_t s, bool fp32, const SegPcgArgs& a, int k) {
  if (a.bcol16) {
    if (fp32)
      hipLaunchKernelGGL((seg_k1_kernel<float, L, true, CG>), dim3(a.W), dim3(K1_BS), 0, s, a, k);
    else
      hipLaunchKernelGGL((seg_k1_kernel<double, L, true, CG>), dim3(a.W), dim3(K1_BS), 0, s, a,
                         k);
  } else {
    if (fp32)
      hipLaunchKernelGGL((seg_k1_kernel<float, L, false, CG>), dim3(a.W), dim3(K1_BS), 0, s, a,
                         k);
    else
      hipLaunchKernelGGL((seg_k1_kernel<double, L, false, CG>), dim3(a.W), dim3(K1_BS), 0, s, a,
                         k);
  }
}

template <int L>
static void seg_k1_l(hipStream_t s, bool fp32, const SegPcgArgs& a, int k) {
  if (a.coarse)
    seg_k1_lc<L, true>(s, fp32, a, k);
  else
    seg_k1_lc<L, false>(s, fp32, a, k);
}

void launch_seg_k1(hipStream_t s, bool fp32, const SegPcgArgs& a, int k) {
  switch (a.lpr) {
    case 4: seg_k1_l<4>(s, fp32, a, k); break;
    case 16: seg_k1_l<16>(s, fp32, a, k); break;
    default: seg_k1_l<8>(s, fp32, a, k); break;
  }
}

void launch_seg_k2(hipStream_t s, bool fp32, const SegPcgArgs& a, int k) {
  if (fp32)
    hipLaunchKernelGGL(seg_k2_kernel<float>, dim3(a.W), dim3(SEG_K2_BS), 0, s, a, k);
  else
    hipLaunchKernelGGL(seg_k2_kernel<double>, dim3(a.W), dim3(SEG_K2_BS), 0, s, a, k);
}

void launch_seg_status(hipStream_t s, const SegPcgArgs& a, int k_next, int maxit) {
  hipLaunchKernelGGL(seg_status_kernel, dim3(a.K), dim3(256), 0, s, a, k_next, maxit);
}

}  // namespace pgn
