// coarse_kernels.hip -- the coarse level of the two-level PCG preconditioner.
//
// M^-1 r = B^-1 r + P Ac^-1 P^T r: B is the 3x3 block-Jacobi part (applied by
// the PCG update launch), P the aggregated rigid-body prolongation described
// with CoarseArgs (kernels.hpp).  Per linear solve try this file builds
// Ac = P^T A P from the assembled BSR3 values (FP64 sums over pieces of at
// most COARSE_PIECE blocks of a coarse block pair, then the pieces in order)
// and inverts it in place by block Gauss-Jordan sweeps (the sweep operator on
// symmetric positive definite matrices: after sweeping every pivot the matrix
// holds -Ac^-1; no pivoting is needed since every pivot is a Schur
// complement of an SPD matrix).  Per PCG iteration it restricts r (four
// workgroups per aggregate), multiplies by the inverse (one row per
// wavefront) and prolongs into z (one node per thread).  Every sum has a fixed order: the results do
// not depend on scheduling.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "pgn_internal.hpp"
#include "reduce.hpp"

namespace pgn {

namespace {

constexpr int TL = COARSE_TILE;  // 32
constexpr int CB_BS = 256;       // coarse assembly workgroup
constexpr int GJ_BS = 256;       // sweep workgroup: 4 entries of a 32 x 32 tile each
constexpr int RS_BS = 256;       // restriction workgroup: a quarter of an aggregate
constexpr int RS_U = 4;          // nodes in flight per thread
constexpr int GV_BS = 256;       // coarse product workgroup: one row per wavefront
constexpr int GV_ROWS = GV_BS / 64;
constexpr int PR_BS = 256;       // prolongation workgroup: one node per thread

// w x r as the 3x3 matrix W(r) acting on w: rows (0, rz, -ry), (-rz, 0, rx),
// (ry, -rx, 0)
__device__ __forceinline__ void rot_block(const double r[3], double W[9]) {
  W[0] = 0.0;   W[1] = r[2];  W[2] = -r[1];
  W[3] = -r[2]; W[4] = 0.0;   W[5] = r[0];
  W[6] = r[1];  W[7] = -r[0]; W[8] = 0.0;
}

// ---- Ac = P^T A P ----------------------------------------------------------------

__global__ __launch_bounds__(256) void coarse_fill_kernel(int npad, int nact, double* A,
                                                          int* ok) {
  const long total = (long)npad * npad;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const long i = e / npad, j = e % npad;
    A[e] = (i == j && i >= nact) ? 1.0 : 0.0;  // padding DoFs: identity
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *ok = 1;
}

// One workgroup per piece (at most COARSE_PIECE listed BSR3 blocks of one
// coarse block pair (a, b), a <= b): the partial 6x6 block
// sum P_i^T A_ij P_j over its blocks, with A_ij's fixed rows / columns masked.
template <class T>
__global__ __launch_bounds__(CB_BS) void coarse_piece_kernel(CoarseArgs c, CoarseBuild b) {
  __shared__ double red[CB_BS / 64][36];
  const int pc = blockIdx.x;
  const T* __restrict__ bval = reinterpret_cast<const T*>(b.bval);
  const T* __restrict__ rel = reinterpret_cast<const T*>(c.rel);
  double acc[36];
#pragma unroll
  for (int e = 0; e < 36; ++e) acc[e] = 0.0;
  const int lo = b.piece_lo[pc], hi = b.piece_lo[pc + 1];
  if (b.live && b.live[b.agg_inst[c.agg_of[b.brow[b.cb_fine[lo]]]]] == 0) return;
  for (int t = lo + threadIdx.x; t < hi; t += CB_BS) {
    const int f = b.cb_fine[t];
    const int i = b.brow[f], j = b.bcol[f];
    const unsigned fi = c.fixmask[i], fj = c.fixmask[j];
    double B[9];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q)
        B[3 * p + q] = (((fi >> p) | (fj >> q)) & 1u) ? 0.0 : (double)bval[9L * f + 3 * p + q];
    const double ri[3] = {(double)rel[3L * i], (double)rel[3L * i + 1], (double)rel[3L * i + 2]};
    const double rj[3] = {(double)rel[3L * j], (double)rel[3L * j + 1], (double)rel[3L * j + 2]};
    double Wi[9], Wj[9];
    rot_block(ri, Wi);
    rot_block(rj, Wj);
    double BW[9];  // B W_j
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q)
        BW[3 * p + q] = B[3 * p] * Wj[q] + B[3 * p + 1] * Wj[3 + q] + B[3 * p + 2] * Wj[6 + q];
    // rows 0-2: [B, B W_j]; rows 3-5: [W_i^T B, W_i^T B W_j]
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        acc[6 * p + q] += B[3 * p + q];
        acc[6 * p + 3 + q] += BW[3 * p + q];
        acc[6 * (3 + p) + q] +=
            Wi[p] * B[q] + Wi[3 + p] * B[3 + q] + Wi[6 + p] * B[6 + q];
        acc[6 * (3 + p) + 3 + q] +=
            Wi[p] * BW[q] + Wi[3 + p] * BW[3 + q] + Wi[6 + p] * BW[6 + q];
      }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int e = 0; e < 36; ++e) {
    const double v = wave_sum(acc[e]);
    if (lane == 0) red[w][e] = v;
  }
  __syncthreads();
  if (threadIdx.x < 36) {
    const int e = threadIdx.x;
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < CB_BS / 64; ++k) v += red[k][e];
    b.piece_sum[36L * pc + e] = v;
  }
}

// One thread per (pair, entry): the pair's pieces summed in order, written
// at (a, b) and transposed at (b, a); inactive coarse columns get zero rows
// and columns and a unit diagonal.
__global__ __launch_bounds__(256) void coarse_pairs_kernel(CoarseArgs c, CoarseBuild b) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= 36L * b.ncb) return;
  const int cb = (int)(t / 36), e = (int)(t % 36), p = e / 6, q = e % 6;
  const int ca = b.cb_a[cb], cbb = b.cb_b[cb];
  double v = 0.0;
  for (int k = b.pair_piece0[cb]; k < b.pair_piece0[cb + 1]; ++k) v += b.piece_sum[36L * k + e];
  const bool act_p = (c.cmask[ca] >> p) & 1u, act_q = (c.cmask[cbb] >> q) & 1u;
  if (!act_p || !act_q) v = (ca == cbb && p == q) ? 1.0 : 0.0;
  const long rI = 6L * ca + p, cJ = 6L * cbb + q;
  b.A0[rI * c.npad + cJ] = v;
  if (ca != cbb) b.A0[cJ * c.npad + rI] = v;
}

// ---- block Gauss-Jordan sweeps -----------------------------------------------------

// Sweep of the 32x32 matrix in Ma (LDS, row-major) over all its pivots by
// the whole workgroup, 4 entries per thread: pivot p reads one of the
// buffers Ma, Mb and writes the other (one barrier per pivot); after the 32
// pivots Ma holds -M^-1.  d0 (LDS, 32) receives the starting diagonal.
// Returns false when a pivot is not positive relative to its starting
// diagonal value (all threads agree).
__device__ bool sweep_tile(double* Ma, double* Mb, double* d0) {
  const int t = threadIdx.x;
  if (t < TL) d0[t] = Ma[t * TL + t];
  __syncthreads();
  bool good = true;
  for (int p = 0; p < TL; ++p) {
    const double* src = (p & 1) ? Mb : Ma;
    double* dst = (p & 1) ? Ma : Mb;
    double d = src[p * TL + p];
    if (!(d > 1e-13 * fabs(d0[p]) && d > 0.0)) {
      good = false;
      d = 1.0;  // keeps the remaining arithmetic finite; the result is flagged
    }
    const double inv = 1.0 / d;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = t + GJ_BS * q, i = e / TL, j = e % TL;
      const double mij = src[e], mip = src[i * TL + p], mpj = src[p * TL + j];
      double v;
      if (i == p && j == p) v = -inv;
      else if (i == p) v = mpj * inv;
      else if (j == p) v = mip * inv;
      else v = mij - mip * mpj * inv;
      dst[e] = v;
    }
    __syncthreads();
  }
  return good;
}

__device__ __forceinline__ void load_tile(const double* __restrict__ A, int npad, int I, int J,
                                          double* T) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = threadIdx.x + GJ_BS * q, i = e / TL, j = e % TL;
    T[e] = A[(long)(I * TL + i) * npad + J * TL + j];
  }
}

// C = X Y for 32x32 tiles in LDS (4 entries per thread: row t / 8, columns
// t % 8 + 8 q); C may not alias X or Y
__device__ __forceinline__ void tile_mul(const double* X, const double* Y, double out[4]) {
  const int r = threadIdx.x >> 3, c0 = threadIdx.x & 7;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll 8
  for (int m = 0; m < TL; ++m) {
    const double x = X[r * TL + m];
    a0 += x * Y[m * TL + c0];
    a1 += x * Y[m * TL + c0 + 8];
    a2 += x * Y[m * TL + c0 + 16];
    a3 += x * Y[m * TL + c0 + 24];
  }
  out[0] = a0;
  out[1] = a1;
  out[2] = a2;
  out[3] = a3;
}

// D0 = inverse of the first pivot tile
__global__ __launch_bounds__(GJ_BS) void coarse_pivot0_kernel(int npad, const double* A,
                                                              double* D, int* ok) {
  __shared__ double M[TL * TL];
  __shared__ double M2[TL * TL];
  __shared__ double d0[TL];
  load_tile(A, npad, 0, 0, M);
  __syncthreads();
  const bool good = sweep_tile(M, M2, d0);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = threadIdx.x + GJ_BS * q;
    D[e] = -M[e];
  }
  if (!good && threadIdx.x == 0) *ok = 0;
}

// Sweep of pivot tile k for tile (I, J) = (blockIdx.y, blockIdx.x), src ->
// dst, with Dk = A_kk^-1:
//   (k, k): -Dk;  (k, J): Dk A_kJ;  (I, k): A_Ik Dk;
//   else:   A_IJ - (A_Ik Dk) A_kJ.
// The workgroup of tile (k+1, k+1) then inverts its new tile: the next pivot.
__global__ __launch_bounds__(GJ_BS) void coarse_sweep_kernel(int npad, int k,
                                                             const double* __restrict__ src,
                                                             double* __restrict__ dst,
                                                             const double* __restrict__ Dk,
                                                             double* __restrict__ Dnext, int* ok) {
  __shared__ double sD[TL * TL];
  __shared__ double sX[TL * TL];
  __shared__ double sY[TL * TL];
  __shared__ double sC[TL * TL];
  __shared__ double d0[TL];
  const int I = blockIdx.y, J = blockIdx.x;
  const int r = threadIdx.x >> 3, c0 = threadIdx.x & 7;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = threadIdx.x + GJ_BS * q;
    sD[e] = Dk[e];
  }
  double out[4];
  if (I == k && J == k) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) out[q] = -sD[r * TL + c0 + 8 * q];
  } else if (I == k) {
    load_tile(src, npad, k, J, sY);
    __syncthreads();
    tile_mul(sD, sY, out);
  } else if (J == k) {
    load_tile(src, npad, I, k, sX);
    __syncthreads();
    tile_mul(sX, sD, out);
  } else {
    load_tile(src, npad, I, k, sX);
    load_tile(src, npad, k, J, sY);
    __syncthreads();
    tile_mul(sX, sD, out);
#pragma unroll
    for (int q = 0; q < 4; ++q) sC[r * TL + c0 + 8 * q] = out[q];
    __syncthreads();
    tile_mul(sC, sY, out);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = r, j = c0 + 8 * q;
      out[q] = src[(long)(I * TL + i) * npad + J * TL + j] - out[q];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i = r, j = c0 + 8 * q;
    dst[(long)(I * TL + i) * npad + J * TL + j] = out[q];
  }
  if (I == k + 1 && J == k + 1) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) sX[r * TL + c0 + 8 * q] = out[q];
    __syncthreads();
    const bool good = sweep_tile(sX, sC, d0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = threadIdx.x + GJ_BS * q;
      Dnext[e] = -sX[e];
    }
    if (!good && threadIdx.x == 0) *ok = 0;
  }
}

// R := (R + R^T) / 2 (the sweeps' rounding leaves the inverse symmetric only
// to rounding; the preconditioner must be exactly symmetric)
__global__ __launch_bounds__(256) void coarse_symmetrize_kernel(int npad, double* R) {
  const long total = (long)npad * npad;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const long i = e / npad, j = e % npad;
    if (i < j) {
      const double v = 0.5 * (R[e] + R[j * npad + i]);
      R[e] = v;
      R[j * npad + i] = v;
    }
  }
}

// ---- per PCG iteration ------------------------------------------------------------

// Part s (blockIdx.y) of P^T r for aggregate a (blockIdx.x): translations sum
// F_i r_i, rotations sum rel_i x F_i r_i, over the s-th quarter of the
// aggregate's nodes in order (RS_U nodes per thread in flight), into
// rc[s npad + 6 a + e]; dropped columns get zero.
template <class T>
__global__ __launch_bounds__(RS_BS) void coarse_restrict_kernel(PcgArgs A) {
  __shared__ double red[RS_BS / 64][6];
  if (A.sc->done) return;
  const CoarseArgs& c = A.cg;
  const T* __restrict__ r = reinterpret_cast<const T*>(A.r);
  const T* __restrict__ rel = reinterpret_cast<const T*>(c.rel);
  const int a = blockIdx.x, part = blockIdx.y;
  const int a0 = c.agg_ptr[a], len = c.agg_ptr[a + 1] - a0;
  const int lo = a0 + (int)((long)len * part / COARSE_SPLIT);
  const int hi = a0 + (int)((long)len * (part + 1) / COARSE_SPLIT);
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int t0 = lo + threadIdx.x; t0 < hi; t0 += RS_BS * RS_U) {
    int idx[RS_U];
#pragma unroll
    for (int u = 0; u < RS_U; ++u) {
      const int t = t0 + u * RS_BS;
      idx[u] = t < hi ? c.agg_nodes[t] : -1;
    }
    T rv[RS_U][3], rl[RS_U][3];
    unsigned fm[RS_U];
#pragma unroll
    for (int u = 0; u < RS_U; ++u) {
      const long i = idx[u] < 0 ? 0 : idx[u];
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        rv[u][p] = r[3 * i + p];
        rl[u][p] = rel[3 * i + p];
      }
      fm[u] = idx[u] < 0 ? 7u : (unsigned)c.fixmask[i];
    }
#pragma unroll
    for (int u = 0; u < RS_U; ++u) {
      double v[3];
#pragma unroll
      for (int p = 0; p < 3; ++p) v[p] = ((fm[u] >> p) & 1u) ? 0.0 : (double)rv[u][p];
      const double x = rl[u][0], y = rl[u][1], z = rl[u][2];
      acc[0] += v[0];
      acc[1] += v[1];
      acc[2] += v[2];
      acc[3] += y * v[2] - z * v[1];
      acc[4] += z * v[0] - x * v[2];
      acc[5] += x * v[1] - y * v[0];
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const double v = wave_sum(acc[e]);
    if (lane == 0) red[w][e] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int e = threadIdx.x;
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < RS_BS / 64; ++k) v += red[k][e];
    if (!((c.cmask[a] >> e) & 1u)) v = 0.0;
    c.rc[(long)part * c.npad + 6L * a + e] = v;
  }
}

// rc_j: its COARSE_SPLIT parts summed in order
__device__ __forceinline__ double rc_at(const double* __restrict__ rc, int npad, int j) {
  double v = rc[j];
#pragma unroll
  for (int s = 1; s < COARSE_SPLIT; ++s) v += rc[(long)s * npad + j];
  return v;
}

// uc = Ac^-1 rc (R holds -Ac^-1), one row per wavefront; one partial of
// rc.uc per workgroup.  uc = 0 when the sweeps failed.
__global__ __launch_bounds__(GV_BS) void coarse_gemv_kernel(PcgArgs A) {
  __shared__ double red[GV_BS / 64];
  if (A.sc->done) return;
  const CoarseArgs& c = A.cg;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row = blockIdx.x * GV_ROWS + w;
  const double* __restrict__ Rr = c.R + (long)row * c.npad;
  const double* __restrict__ rc = c.rc;
  const bool good = *c.ok != 0;
  double v0 = 0.0, v1 = 0.0;
  int j = lane;
  for (; j + 64 < c.npad; j += 128) {
    v0 += Rr[j] * rc_at(rc, c.npad, j);
    v1 += Rr[j + 64] * rc_at(rc, c.npad, j + 64);
  }
  if (j < c.npad) v0 += Rr[j] * rc_at(rc, c.npad, j);
  const double v = wave_sum(v0 + v1);
  double dl = 0.0;
  if (lane == 0) {
    const double u = good ? -v : 0.0;
    c.uc[row] = u;
    dl = rc_at(rc, c.npad, row) * u;
  }
  const double sdl = block_sum<GV_BS>(dl, red);
  if (threadIdx.x == 0) c.cpart[blockIdx.x] = sdl;
}

// z_i += F_i (t_a + w_a x rel_i), one node per thread; block 0 also sums
// the rc.uc partials
template <class T>
__global__ __launch_bounds__(PR_BS) void coarse_prolong_kernel(PcgArgs A) {
  __shared__ double red[PR_BS / 64];
  if (A.sc->done) return;
  const CoarseArgs& c = A.cg;
  if (blockIdx.x == 0) {
    double v = 0.0;
    for (int g = threadIdx.x; g < c.gemv_g; g += PR_BS) v += c.cpart[g];
    v = block_sum<PR_BS>(v, red);
    if (threadIdx.x == 0) *c.cdot = v;
  }
  const long i = (long)blockIdx.x * PR_BS + threadIdx.x;
  if (i >= A.n) return;
  T* __restrict__ z = reinterpret_cast<T*>(A.z);
  const T* __restrict__ rel = reinterpret_cast<const T*>(c.rel);
  const int a = c.agg_of[i];
  const unsigned fm = c.fixmask[i];
  const T z0 = z[3 * i], z1 = z[3 * i + 1], z2 = z[3 * i + 2];
  const double x = rel[3 * i], y = rel[3 * i + 1], zz = rel[3 * i + 2];
  const double* u = c.uc + 6L * a;
  const double d0 = u[0] + u[4] * zz - u[5] * y;
  const double d1 = u[1] + u[5] * x - u[3] * zz;
  const double d2 = u[2] + u[3] * y - u[4] * x;
  if (!(fm & 1u)) z[3 * i] = (T)((double)z0 + d0);
  if (!(fm & 2u)) z[3 * i + 1] = (T)((double)z1 + d1);
  if (!(fm & 4u)) z[3 * i + 2] = (T)((double)z2 + d2);
}

// ---- segmented (one coarse operator per instance of a batch) ------------------

constexpr int SG_BS = 512;  // sweep / fill workgroup: one instance each
constexpr int SG_MAXDOF = 6 * SEG_COARSE_MAXAGG;

__device__ __forceinline__ int seg_k_abs(const SegPcgArgs& A, int k) {
  return A.kctl ? A.kctl[0] + k : k;
}

// one workgroup per instance: its dense operator zeroed, ok = 1
__global__ __launch_bounds__(SG_BS) void seg_coarse_fill_kernel(SegPcgArgs A, double* M, int* ok) {
  const int sg = blockIdx.x;
  if (A.seg_G[sg] == 0) return;
  const SegCoarseArgs& c = A.cg;
  if (c.use && !c.use[sg]) {  // block Jacobi for this instance: uc = 0
    if (threadIdx.x == 0) ok[sg] = 0;
    return;
  }
  const int nc = 6 * (c.inst_aoff[sg + 1] - c.inst_aoff[sg]);
  double* Mk = M + c.inst_moff[sg];
  for (int e = threadIdx.x; e < nc * nc; e += SG_BS) Mk[e] = 0.0;
  if (threadIdx.x == 0) ok[sg] = 1;
}

// as coarse_pairs_kernel, into the instance's own operator
__global__ __launch_bounds__(256) void seg_coarse_pairs_kernel(SegPcgArgs A, CoarseBuild b,
                                                               double* M) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= 36L * b.ncb) return;
  const SegCoarseArgs& c = A.cg;
  const int cb = (int)(t / 36), e = (int)(t % 36), p = e / 6, q = e % 6;
  const int ca = b.cb_a[cb], cbb = b.cb_b[cb];
  const int sg = c.agg_inst[ca];
  if (A.seg_G[sg] == 0 || (c.use && !c.use[sg])) return;
  double v = 0.0;
  for (int k = b.pair_piece0[cb]; k < b.pair_piece0[cb + 1]; ++k) v += b.piece_sum[36L * k + e];
  const bool act_p = (c.cmask[ca] >> p) & 1u, act_q = (c.cmask[cbb] >> q) & 1u;
  if (!act_p || !act_q) v = (ca == cbb && p == q) ? 1.0 : 0.0;
  const int a0 = c.inst_aoff[sg], nc = 6 * (c.inst_aoff[sg + 1] - a0);
  const long rI = 6L * (ca - a0) + p, cJ = 6L * (cbb - a0) + q;
  double* Mk = M + c.inst_moff[sg];
  Mk[rI * nc + cJ] = v;
  if (ca != cbb) Mk[cJ * nc + rI] = v;
}

// One workgroup per instance: Gauss-Jordan sweep of its nc x nc operator in
// place (pivot row and column staged in LDS, two barriers per pivot; the
// entries stay in the workgroup's own cache lines), then symmetrized.  After
// the nc pivots the storage holds -Ac^-1; a pivot that is not positive
// relative to its starting diagonal clears ok.
__global__ __launch_bounds__(SG_BS) void seg_coarse_sweep_kernel(SegPcgArgs A, double* M,
                                                                 int* ok) {
  __shared__ double rowp[SG_MAXDOF], colp[SG_MAXDOF], dst0[SG_MAXDOF];
  const int sg = blockIdx.x;
  if (A.seg_G[sg] == 0) return;
  const SegCoarseArgs& c = A.cg;
  if (c.use && !c.use[sg]) return;
  const int nc = 6 * (c.inst_aoff[sg + 1] - c.inst_aoff[sg]);
  double* Mk = M + c.inst_moff[sg];
  const int t = threadIdx.x;
  for (int e = t; e < nc; e += SG_BS) dst0[e] = Mk[e * nc + e];
  bool good = true;
  for (int p = 0; p < nc; ++p) {
    __syncthreads();
    for (int e = t; e < nc; e += SG_BS) {
      rowp[e] = Mk[p * nc + e];
      colp[e] = Mk[e * nc + p];
    }
    __syncthreads();
    double d = rowp[p];
    if (!(d > 1e-13 * fabs(dst0[p]) && d > 0.0)) {
      good = false;
      d = 1.0;  // keeps the remaining arithmetic finite; the result is flagged
    }
    const double inv = 1.0 / d;
    for (int e = t; e < nc * nc; e += SG_BS) {
      const int i = e / nc, j = e - i * nc;
      double v;
      if (i == p && j == p) v = -inv;
      else if (i == p) v = rowp[j] * inv;
      else if (j == p) v = colp[i] * inv;
      else v = Mk[e] - colp[i] * rowp[j] * inv;
      Mk[e] = v;
    }
  }
  __syncthreads();
  for (int e = t; e < nc * nc; e += SG_BS) {
    const int i = e / nc, j = e - i * nc;
    if (i < j) {
      const double v = 0.5 * (Mk[e] + Mk[j * nc + i]);
      Mk[e] = v;
      Mk[j * nc + i] = v;
    }
  }
  if (!good && t == 0) ok[sg] = 0;
}

constexpr int SA_BS = 1024;  // segmented coarse product: one instance per workgroup

// P^T r for aggregate a (blockIdx.x), its nodes in order, RS_U per thread in
// flight, into rc[6 a .. 6 a + 5]
template <class T>
__global__ __launch_bounds__(RS_BS) void seg_coarse_restrict_kernel(SegPcgArgs A) {
  __shared__ double red[RS_BS / 64][6];
  const SegCoarseArgs& c = A.cg;
  const int a = blockIdx.x, sg = c.agg_inst[a];
  if (A.seg_G[sg] == 0 || A.sc[sg].done || (c.use && !c.use[sg])) return;
  const T* __restrict__ r = reinterpret_cast<const T*>(A.r);
  const T* __restrict__ rel = reinterpret_cast<const T*>(c.rel);
  const int lo = c.agg_ptr[a], hi = c.agg_ptr[a + 1];
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int t0 = lo + threadIdx.x; t0 < hi; t0 += RS_BS * RS_U) {
    int idx[RS_U];
#pragma unroll
    for (int u = 0; u < RS_U; ++u) {
      const int t = t0 + u * RS_BS;
      idx[u] = t < hi ? c.agg_nodes[t] : -1;
    }
    T rv[RS_U][3], rl[RS_U][3];
    unsigned fm[RS_U];
#pragma unroll
    for (int u = 0; u < RS_U; ++u) {
      const long i = idx[u] < 0 ? 0 : idx[u];
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        rv[u][p] = r[3 * i + p];
        rl[u][p] = rel[3 * i + p];
      }
      fm[u] = idx[u] < 0 ? 7u : (unsigned)c.fixmask[i];
    }
#pragma unroll
    for (int u = 0; u < RS_U; ++u) {
      double v[3];
#pragma unroll
      for (int p = 0; p < 3; ++p) v[p] = ((fm[u] >> p) & 1u) ? 0.0 : (double)rv[u][p];
      const double x = rl[u][0], y = rl[u][1], z = rl[u][2];
      acc[0] += v[0];
      acc[1] += v[1];
      acc[2] += v[2];
      acc[3] += y * v[2] - z * v[1];
      acc[4] += z * v[0] - x * v[2];
      acc[5] += x * v[1] - y * v[0];
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const double v = wave_sum(acc[e]);
    if (lane == 0) red[w][e] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int e = threadIdx.x;
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < RS_BS / 64; ++q) v += red[q][e];
    if (!((c.cmask[a] >> e) & 1u)) v = 0.0;
    c.rc[6L * a + e] = v;
  }
}

// uc = Ac^-1 rc for instance blockIdx.x: R is symmetric, so thread (q, i)
// sums column i over the q-th of SA_BS / nc column ranges (loads coalesced
// across i), the ranges are added in order, and rc.uc goes into
// cdot[(k & 1) K + instance].  uc = 0 when the instance's sweep failed.
__global__ __launch_bounds__(SA_BS) void seg_coarse_gemv_kernel(SegPcgArgs A, int k) {
  static_assert(SG_MAXDOF <= SA_BS, "one coarse row per thread at least");
  __shared__ double src[SG_MAXDOF];
  __shared__ double part[SA_BS];
  __shared__ double red1[SA_BS / 64];
  const int sg = blockIdx.x;
  if (A.seg_G[sg] == 0 || A.sc[sg].done) return;
  k = seg_k_abs(A, k);
  const SegCoarseArgs& c = A.cg;
  const int t = threadIdx.x;
  const int a0 = c.inst_aoff[sg], nc = 6 * (c.inst_aoff[sg + 1] - a0);
  if (nc == 0) return;
  if (t < nc) src[t] = c.rc[6L * a0 + t];
  __syncthreads();
  const double* __restrict__ Mk = c.R + c.inst_moff[sg];
  const int P = SA_BS / nc;
  if (t < P * nc) {
    const int i = t % nc, q = t / nc;
    const int j0 = q * nc / P, j1 = (q + 1) * nc / P;
    double v0 = 0.0, v1 = 0.0;
    int j = j0;
    for (; j + 2 <= j1; j += 2) {
      v0 += Mk[(long)j * nc + i] * src[j];
      v1 += Mk[(long)(j + 1) * nc + i] * src[j + 1];
    }
    if (j < j1) v0 += Mk[(long)j * nc + i] * src[j];
    part[t] = v0 + v1;
  }
  __syncthreads();
  double dl = 0.0;
  const bool good = c.ok[sg] != 0;  // 0 also for instances without the coarse level
  if (t < nc) {
    double v = 0.0;
    for (int q = 0; q < P; ++q) v += part[q * nc + t];
    const double u = good ? -v : 0.0;
    c.uc[6L * a0 + t] = u;
    dl = good ? src[t] * u : 0.0;  // rc may be stale where the level is off
  }
  const double s = block_sum<SA_BS>(dl, red1);
  if (t == 0) c.cdot[(long)(k & 1) * A.K + sg] = s;
}

}  // namespace

double* launch_coarse_operator(hipStream_t s, bool fp32, const CoarseArgs& c,
                               const CoarseBuild& b, bool assemble_only) {
  const int nt = c.npad / TL;
  const int fg = (int)std::min<long>(((long)c.npad * c.npad + 255) / 256, 1024);
  hipLaunchKernelGGL(coarse_fill_kernel, dim3(fg), dim3(256), 0, s, c.npad, 6 * c.na, b.A0,
                     b.ok);
  if (b.npiece > 0) {
    if (fp32)
      hipLaunchKernelGGL(coarse_piece_kernel<float>, dim3(b.npiece), dim3(CB_BS), 0, s, c, b);
    else
      hipLaunchKernelGGL(coarse_piece_kernel<double>, dim3(b.npiece), dim3(CB_BS), 0, s, c, b);
    const int pg = (int)((36L * b.ncb + 255) / 256);
    hipLaunchKernelGGL(coarse_pairs_kernel, dim3(pg), dim3(256), 0, s, c, b);
  }
  if (assemble_only) return b.A0;
  hipLaunchKernelGGL(coarse_pivot0_kernel, dim3(1), dim3(GJ_BS), 0, s, c.npad, b.A0, b.D, b.ok);
  double* bufs[2] = {b.A0, b.A1};
  for (int k = 0; k < nt; ++k)
    hipLaunchKernelGGL(coarse_sweep_kernel, dim3(nt, nt), dim3(GJ_BS), 0, s, c.npad, k,
                       bufs[k & 1], bufs[(k + 1) & 1], b.D + (k & 1) * TL * TL,
                       b.D + ((k + 1) & 1) * TL * TL, b.ok);
  double* R = bufs[nt & 1];
  hipLaunchKernelGGL(coarse_symmetrize_kernel, dim3(fg), dim3(256), 0, s, c.npad, R);
  return R;
}

void launch_coarse_apply(hipStream_t s, bool fp32, const PcgArgs& a) {
  const CoarseArgs& c = a.cg;
  const int pg = (a.n + PR_BS - 1) / PR_BS;
  if (fp32) {
    hipLaunchKernelGGL(coarse_restrict_kernel<float>, dim3(c.na, COARSE_SPLIT), dim3(RS_BS), 0,
                       s, a);
    hipLaunchKernelGGL(coarse_gemv_kernel, dim3(c.gemv_g), dim3(GV_BS), 0, s, a);
    hipLaunchKernelGGL(coarse_prolong_kernel<float>, dim3(pg), dim3(PR_BS), 0, s, a);
  } else {
    hipLaunchKernelGGL(coarse_restrict_kernel<double>, dim3(c.na, COARSE_SPLIT), dim3(RS_BS),
                       0, s, a);
    hipLaunchKernelGGL(coarse_gemv_kernel, dim3(c.gemv_g), dim3(GV_BS), 0, s, a);
    hipLaunchKernelGGL(coarse_prolong_kernel<double>, dim3(pg), dim3(PR_BS), 0, s, a);
  }
}

int coarse_gemv_grid(int npad) { return npad / GV_ROWS; }

void launch_seg_coarse_operator(hipStream_t s, bool fp32, const SegPcgArgs& a,
                                const CoarseBuild& b, double* A, int* ok) {
  hipLaunchKernelGGL(seg_coarse_fill_kernel, dim3(a.K), dim3(SG_BS), 0, s, a, A, ok);
  if (b.npiece > 0) {
    CoarseArgs c;  // what the piece kernel reads
    c.rel = a.cg.rel;
    c.fixmask = a.cg.fixmask;
    c.agg_of = a.cg.agg_of;
    CoarseBuild bs = b;
    bs.agg_inst = a.cg.agg_inst;
    bs.live = a.seg_G;
    if (fp32)
      hipLaunchKernelGGL(coarse_piece_kernel<float>, dim3(b.npiece), dim3(CB_BS), 0, s, c, bs);
    else
      hipLaunchKernelGGL(coarse_piece_kernel<double>, dim3(b.npiece), dim3(CB_BS), 0, s, c, bs);
    const int pg = (int)((36L * b.ncb + 255) / 256);
    hipLaunchKernelGGL(seg_coarse_pairs_kernel, dim3(pg), dim3(256), 0, s, a, bs, A);
  }
  hipLaunchKernelGGL(seg_coarse_sweep_kernel, dim3(a.K), dim3(SG_BS), 0, s, a, A, ok);
}

void launch_seg_coarse_apply(hipStream_t s, bool fp32, const SegPcgArgs& a, int k) {
  if (fp32)
    hipLaunchKernelGGL(seg_coarse_restrict_kernel<float>, dim3(a.cg.NA), dim3(RS_BS), 0, s, a);
  else
    hipLaunchKernelGGL(seg_coarse_restrict_kernel<double>, dim3(a.cg.NA), dim3(RS_BS), 0, s, a);
  hipLaunchKernelGGL(seg_coarse_gemv_kernel, dim3(a.K), dim3(SA_BS), 0, s, a, k);
}


}  // namespace pgn
