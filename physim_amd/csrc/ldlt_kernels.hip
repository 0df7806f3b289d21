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
    __syncthreads();
#pragma unroll
    for (int q = 0; q < LDLT_Q; ++q) {
      const int ti = pk[q] >> 8, tj = pk[q] & 255;
      if (pk[q] < 0 || ti >= m) continue;
      int si = sk + 1 + ti;
      if (si >= W) si -= W;
      win[si * W + (ti - tj)] -= lv[ti] * d * lv[tj];
    }
    __syncthreads();
    // row k is done; its slot takes row k + W, which no earlier column touches
    const long rn = (long)k + W;
    if (rn < N)
      for (int t = tid; t < W; t += LDLT_FBS) win[sk * W + t] = band[rn * W + t];
    sk = sk + 1 == W ? 0 : sk + 1;
    __syncthreads();
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
  constexpr int S = (LDLT_MAXW + 63) / 64;  // band entries per lane and row
  // forward: y_I = b_I - sum_{d=1..min(bw,I)} L(I, I-d) y_{I-d}; ring[I mod W] = y_I
  double cur[S];
#pragma unroll
  for (int q = 0; q < S; ++q) {
    cur[q] = 0.0;  // row 0 has no subdiagonal entries
  }
  int iw = 0;  // I mod W
  for (int I = 0; I < N; ++I) {
    double nxt[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {  // the next row's band entries, loaded ahead
      const int d = 1 + lane + 64 * q;
      nxt[q] = (I + 1 < N && d <= bw && d <= I + 1) ? band[(long)(I + 1) * W + d] : 0.0;
    }
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const int d = 1 + lane + 64 * q;
      if (d <= bw && d <= I) {
        int r = iw - d;
        if (r < 0) r += W;
        acc += cur[q] * ring[r];
      }
    }
    acc = wave_sum(acc);
    if (lane == 0) {
      const double yi = b[ldlt_dof(A, P, I)] - acc;
      ring[iw] = yi;
      y[I] = yi;
    }
#pragma unroll
    for (int q = 0; q < S; ++q) cur[q] = nxt[q];
    iw = iw + 1 == W ? 0 : iw + 1;
    __syncthreads();
  }
  // backward, right-looking: x_I = y_I / D_I - pend_I, then row I of L adds
  // L(I, I-d) x_I to the pending sums of rows I-1 .. I-bw (ring slots)
  for (int t = lane; t < W; t += 64) ring[t] = 0.0;
  __syncthreads();
  iw = (N - 1) % W;
  for (int I = N - 1; I >= 0; --I) {
    double row[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const int d = 1 + lane + 64 * q;
      row[q] = (d <= bw && d <= I) ? band[(long)I * W + d] : 0.0;
    }
    const double xi = y[I] / D[I] - ring[iw];
    __syncthreads();  // every lane has read ring[iw] before it is cleared
    if (lane == 0) {
      x[ldlt_dof(A, P, I)] = xi;
      ring[iw] = 0.0;  // the slot now serves row I - W
    }
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const int d = 1 + lane + 64 * q;
      if (d <= bw && d <= I) {
        int r = iw - d;
        if (r < 0) r += W;
        ring[r] += row[q] * xi;
      }
    }
    iw = iw == 0 ? W - 1 : iw - 1;
    __syncthreads();
  }
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
