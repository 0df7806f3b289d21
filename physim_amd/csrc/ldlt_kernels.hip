// ldlt_kernels.hip -- FP64 banded LDL^T factorization and solves: the
// EigenLDLT linear solver (LinearSolver_EigenLDLT.cpp:49-153, Eigen's
// SimplicialLDLT: L D L^T without pivoting, a zero pivot fails the
// factorization) for models whose band order keeps the half-bandwidth
// within LDLT_MAXW - 1 DoFs.
//
// One workgroup factors one problem column by column.  The rows the next
// columns touch, k..k+bw, live in LDS as a ring of W = bw + 1 rows of W
// doubles (row I in slot I mod W, entry (I, I - d) at offset d); column k
// scales its W - 1 subdiagonal entries by 1 / D_k into L (written to the
// band in HBM), applies the rank-1 update to the (bw x bw) triangle below
// it, and the slot of row k takes row k + W from HBM.  The solves are one
// wavefront per problem: the forward substitution keeps the last W values in
// LDS and reads L row-wise; the backward substitution is right-looking, each
// finished x_I adding L(I, I - d) x_I to the pending sums of the W - 1 rows
// above it.

#include <algorithm>

#include "kernels.hpp"
#include "reduce.hpp"

namespace pgn {

namespace {

constexpr int LDLT_FBS = 1024;  // factorization workgroup

// Workgroup barrier that orders LDS only: global loads issued before it stay
// in flight (a full __syncthreads waits for them).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
// The same for a one-wavefront workgroup: LDS operations of a wavefront
// complete in order, so only the compiler's ordering is needed.
__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
}
// triangle pairs (ti >= tj) of a (W-1) x (W-1) lower triangle per thread
constexpr int LDLT_Q = ((LDLT_MAXW - 1) * LDLT_MAXW / 2 + LDLT_FBS - 1) / LDLT_FBS;

__global__ __launch_bounds__(256) void ldlt_zero_kernel(LdltArgs A, const int* __restrict__ list) {
  const LdltProb P = A.prob[list[blockIdx.y]];
  const long len = (long)P.N * (P.bw + 1);
  double* band = A.band + P.band0;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < len;
       e += (long)gridDim.x * blockDim.x)
    band[e] = 0.0;
}

template <class T>
__global__ __launch_bounds__(256) void ldlt_scatter_kernel(LdltArgs A, long nb,
                                                           const int* __restrict__ brow,
                                                           const int* __restrict__ bcol,
                                                           const T* __restrict__ bval,
                                                           const uint8_t* __restrict__ pmask) {
  for (long s = (long)blockIdx.x * blockDim.x + threadIdx.x; s < nb;
       s += (long)gridDim.x * blockDim.x) {
    const int i = brow[s], j = bcol[s];
    const int p = A.nprob[i];
    if (pmask && !pmask[p]) continue;
    const LdltProb P = A.prob[p];
    const int qi = A.perm[i] - P.node0, qj = A.perm[j] - P.node0;
    if (qi < qj) continue;  // the transposed block carries the same entries
    const int W = P.bw + 1;
    double* band = A.band + P.band0;
    const T* B = bval + 9 * s;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int I = 3 * qi + a, J = 3 * qj + c;
        if (I >= J) band[(long)I * W + (I - J)] = (double)B[3 * a + c];
      }
  }
}

__global__ __launch_bounds__(LDLT_FBS) void ldlt_factor_kernel(LdltArgs A,
                                                               const int* __restrict__ list) {
  __shared__ double win[LDLT_MAXW * LDLT_MAXW];
  __shared__ double lv[LDLT_MAXW];
  const int p = list[blockIdx.x];
  const LdltProb P = A.prob[p];
  const int bw = P.bw, W = bw + 1, N = P.N, tid = threadIdx.x;
  double* band = A.band + P.band0;
  double* D = A.D + P.dof0;
  const double sinT = A.sinT[p];
  // this thread's pairs (ti, tj), ti >= tj, of the bw x bw update triangle
  int pk[LDLT_Q];
  const int tri = bw * (bw + 1) / 2;
#pragma unroll
  for (int q = 0; q < LDLT_Q; ++q) {
    const int f = tid + q * LDLT_FBS;
    pk[q] = -1;
    if (f < tri) {
      int ti = (int)((sqrt(8.0 * f + 1.0) - 1.0) * 0.5);
      while (ti * (ti + 1) / 2 > f) --ti;
      while ((ti + 1) * (ti + 2) / 2 <= f) ++ti;
      pk[q] = (ti << 8) | (f - ti * (ti + 1) / 2);
    }
  }
  // rows 0..W-1 into slots 0..W-1 (rows past N: zeros, never used)
  for (int f = tid; f < W * W; f += LDLT_FBS) {
    const int r = f / W;
    win[f] = r < N ? band[(long)f] : 0.0;
  }
  // the row that refills the first retired slot, loaded ahead
  double pre = (tid < W && W < N) ? band[(long)W * W + tid] : 0.0;
  __syncthreads();
  int zero_at = -1, nsing = 0, sk = 0;  // sk = k mod W
  for (int k = 0; k < N; ++k) {
    const double d = win[sk * W];
    if (d == 0.0) {  // the same value in every thread: a uniform exit
      zero_at = k;
      break;
    }
    const int m = min(bw, N - 1 - k);  // rows k+1 .. k+m below the pivot
    for (int t = tid; t < m; t += LDLT_FBS) {
      int si = sk + 1 + t;
      if (si >= W) si -= W;
      const double l = win[si * W + t + 1] / d;
      lv[t] = l;
      band[(long)(k + 1 + t) * W + t + 1] = l;
    }
    if (tid == 0) {
      D[k] = d;
      if (!(fabs(d) >= sinT)) ++nsing;
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < LDLT_Q; ++q) {
      const int ti = pk[q] >> 8, tj = pk[q] & 255;
      if (pk[q] < 0 || ti >= m) continue;
      int si = sk + 1 + ti;
      if (si >= W) si -= W;
      win[si * W + (ti - tj)] -= lv[ti] * d * lv[tj];
    }
    // row k is done (the update above touches rows k+1..k+m only): its slot
    // takes row k + W, which no earlier column touches (its L entries are
    // written from column k + 1 on); row k + W + 1 is loaded for the next
    // column
    const long rn = (long)k + W;
    if (rn < N && tid < W) {
      win[sk * W + tid] = pre;
      pre = rn + 1 < N ? band[(rn + 1) * W + tid] : 0.0;
    }
    sk = sk + 1 == W ? 0 : sk + 1;
    lds_barrier();
  }
  if (tid == 0) {
    A.status[2 * p] = zero_at;
    A.status[2 * p + 1] = nsing;
  }
}

// DoF I of a problem in band order -> index in the node-major model vectors
__device__ __forceinline__ long ldlt_dof(const LdltArgs& A, const LdltProb& P, int I) {
  return 3L * A.iperm[P.node0 + I / 3] + I % 3;
}

constexpr int LDLT_S = (LDLT_MAXW + 63) / 64;  // band entries per lane and row

// Row I's subdiagonal band entries L(I, I - d), d = 1 + lane + 64 q (zero
// beyond the band or the matrix).  The loads are unconditional (clamped to
// entry 0) so that the compiler can count them in flight.
__device__ __forceinline__ void ldlt_load_row(const double* band, int W, int bw, int N, int I,
                                              int lane, double (&v)[LDLT_S]) {
#pragma unroll
  for (int q = 0; q < LDLT_S; ++q) {
    const int d = 1 + lane + 64 * q;
    const bool ok = I >= 0 && I < N && d <= bw && d <= I;
    const double t = band[ok ? (long)I * W + d : 0L];
    v[q] = ok ? t : 0.0;
  }
}

// a[i] for i clamped into [0, n)
__device__ __forceinline__ double ldlt_ld(const double* a, int i, int n) {
  return a[i < 0 ? 0 : (i >= n ? n - 1 : i)];
}

// DPP shift of a double within rows of 16 lanes (zero from outside the row)
template <int CTRL>
__device__ __forceinline__ double dpp_shr(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_value(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}
// Sum over the wavefront in a fixed order, returned in every lane: shifted
// partial sums within each row of 16 lanes (row_shr 1, 2, 4, 8), then the
// four row totals.
__device__ __forceinline__ double wave_total(double v) {
  v += dpp_shr<0x111>(v);
  v += dpp_shr<0x112>(v);
  v += dpp_shr<0x114>(v);
  v += dpp_shr<0x118>(v);
  return (lane_value(v, 15) + lane_value(v, 31)) + (lane_value(v, 47) + lane_value(v, 63));
}

// forward row I: y_I = b_I - sum_d L(I, I-d) y_{I-d} (ring holds the last W y)
__device__ __forceinline__ void ldlt_fwd_row(int I, int& iw, int W, int bw, int lane,
                                             const double (&row)[LDLT_S], double bI,
                                             double* ring, double* y) {
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < LDLT_S; ++q) {
    const int d = 1 + lane + 64 * q;
    if (d <= bw && d <= I) {
      int r = iw - d;
      if (r < 0) r += W;
      acc += row[q] * ring[r];
    }
  }
  acc = wave_total(acc);
  if (lane == 0) {
    const double yi = bI - acc;
    ring[iw] = yi;
    y[I] = yi;
  }
  iw = iw + 1 == W ? 0 : iw + 1;
  wave_lds_order();
}

// backward row I: x_I = y_I / D_I - pend_I; pend_{I-d} += L(I, I-d) x_I
__device__ __forceinline__ void ldlt_bwd_row(int I, int& iw, int W, int bw, int lane,
                                             const double (&row)[LDLT_S], double yI, double dI,
                                             double* ring, double* y) {
  const double xi = yI / dI - ring[iw];
  wave_lds_order();  // every lane has read ring[iw] before it is cleared
  if (lane == 0) {
    y[I] = xi;
    ring[iw] = 0.0;  // the slot now serves row I - W
  }
#pragma unroll
  for (int q = 0; q < LDLT_S; ++q) {
    const int d = 1 + lane + 64 * q;
    if (d <= bw && d <= I) {
      int r = iw - d;
      if (r < 0) r += W;
      ring[r] += row[q] * xi;
    }
  }
  iw = iw == 0 ? W - 1 : iw - 1;
  wave_lds_order();
}

// One wavefront per problem.  LDLT_P rows are in flight: each of LDLT_P
// register sets is reloaded with the row LDLT_P ahead right after its row is
// processed, so a row's loads have LDLT_P - 1 rows of work to arrive.
constexpr int LDLT_P = 8;

__global__ __launch_bounds__(64) void ldlt_solve_kernel(LdltArgs A, const int* __restrict__ list,
                                                        const double* __restrict__ b,
                                                        double* __restrict__ x,
                                                        double* __restrict__ yscr) {
  __shared__ double ring[LDLT_MAXW];
  const LdltProb P = A.prob[list[blockIdx.x]];
  const int bw = P.bw, W = bw + 1, N = P.N, lane = threadIdx.x;
  const double* band = A.band + P.band0;
  const double* D = A.D + P.dof0;
  double* y = yscr + P.dof0;
  // b in band order
  for (int I = lane; I < N; I += 64) y[I] = b[ldlt_dof(A, P, I)];
  __syncthreads();
  double rs[LDLT_P][LDLT_S], vs[LDLT_P], ds[LDLT_P];
  int iw = 0;
#pragma unroll
  for (int u = 0; u < LDLT_P; ++u) {
    ldlt_load_row(band, W, bw, N, u, lane, rs[u]);
    vs[u] = ldlt_ld(y, u, N);
  }
  for (int base = 0; base < N; base += LDLT_P) {
#pragma unroll
    for (int u = 0; u < LDLT_P; ++u) {
      const int I = base + u;
      if (I < N) ldlt_fwd_row(I, iw, W, bw, lane, rs[u], vs[u], ring, y);
      ldlt_load_row(band, W, bw, N, I + LDLT_P, lane, rs[u]);
      vs[u] = ldlt_ld(y, I + LDLT_P, N);
    }
  }
  for (int t = lane; t < W; t += 64) ring[t] = 0.0;
  __syncthreads();
  iw = (N - 1) % W;
#pragma unroll
  for (int u = 0; u < LDLT_P; ++u) {
    ldlt_load_row(band, W, bw, N, N - 1 - u, lane, rs[u]);
    vs[u] = ldlt_ld(y, N - 1 - u, N);
    ds[u] = ldlt_ld(D, N - 1 - u, N);
  }
  for (int base = N - 1; base >= 0; base -= LDLT_P) {
#pragma unroll
    for (int u = 0; u < LDLT_P; ++u) {
      const int I = base - u;
      if (I >= 0) ldlt_bwd_row(I, iw, W, bw, lane, rs[u], vs[u], ds[u], ring, y);
      ldlt_load_row(band, W, bw, N, I - LDLT_P, lane, rs[u]);
      vs[u] = ldlt_ld(y, I - LDLT_P, N);
      ds[u] = ldlt_ld(D, I - LDLT_P, N);
    }
  }
  __syncthreads();
  for (int I = lane; I < N; I += 64) x[ldlt_dof(A, P, I)] = y[I];
}

}  // namespace

void launch_ldlt_gather(hipStream_t s, bool fp32, const LdltArgs& a, const int* list,
                        int nlist, long max_band, long nb, const int* brow, const int* bcol,
                        const void* bval, const uint8_t* pmask) {
  if (nlist <= 0) return;
  const int gx = (int)std::min<long>(std::max<long>((max_band + 255) / 256, 1), 2048);
  hipLaunchKernelGGL(ldlt_zero_kernel, dim3(gx, nlist), dim3(256), 0, s, a, list);
  if (nb <= 0) return;
  const int g = (int)std::min<long>((nb + 255) / 256, 8192);
  if (fp32)
    hipLaunchKernelGGL(ldlt_scatter_kernel<float>, dim3(g), dim3(256), 0, s, a, nb, brow, bcol,
                       reinterpret_cast<const float*>(bval), pmask);
  else
    hipLaunchKernelGGL(ldlt_scatter_kernel<double>, dim3(g), dim3(256), 0, s, a, nb, brow, bcol,
                       reinterpret_cast<const double*>(bval), pmask);
}

void launch_ldlt_factor(hipStream_t s, const LdltArgs& a, const int* list, int nlist) {
  if (nlist <= 0) return;
  hipLaunchKernelGGL(ldlt_factor_kernel, dim3(nlist), dim3(LDLT_FBS), 0, s, a, list);
}

void launch_ldlt_solve(hipStream_t s, const LdltArgs& a, const int* list, int nlist,
                       const double* b, double* x, double* y) {
  if (nlist <= 0) return;
  hipLaunchKernelGGL(ldlt_solve_kernel, dim3(nlist), dim3(64), 0, s, a, list, b, x, y);
}

}  // namespace pgn
