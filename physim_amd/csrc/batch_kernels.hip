// batch_kernels.hip -- per-instance reductions and updates for a fused batch
// of independent solves.
//
// The batch is one device model whose node and element sets are the
// disjoint union of its instances'; element, gather, assembly and PCG
// kernels run once over the whole batch.  These kernels give each instance
// its own Newton state on top: segmented (one workgroup per instance,
// fixed-order) energy / gradient-norm reductions and masked per-instance
// vector updates.  Instances own contiguous node ranges and, per element
// class, contiguous element ranges.
#include "kernels.hpp"
#include "reduce.hpp"

namespace pgn {

constexpr int BR_BLOCK = 256;

__global__ __launch_bounds__(BR_BLOCK) void batch_norm2_kernel(BatchArgs A,
                                                               const double* __restrict__ v,
                                                               double* __restrict__ out) {
  __shared__ double red[BR_BLOCK / 64];
  const int k = blockIdx.x;
  const long d0 = 3 * A.noff[k], d1 = 3 * A.noff[k + 1];
  double s = 0.0;
  for (long d = d0 + threadIdx.x; d < d1; d += BR_BLOCK) s += v[d] * v[d];
  s = block_sum<BR_BLOCK>(s, red);
  if (threadIdx.x == 0) out[k] = s;
}

__global__ __launch_bounds__(256) void batch_step_kernel(long nd, const int* __restrict__ inst,
                                                         const double* __restrict__ fac,
                                                         const double* __restrict__ xsave,
                                                         const double* __restrict__ dx,
                                                         double* __restrict__ x) {
  const long d = blockIdx.x * 256L + threadIdx.x;
  if (d >= nd) return;
  const double f = fac[inst[d / 3]];
  if (f >= 0.0) x[d] = xsave[d] + f * dx[d];
}

__global__ __launch_bounds__(256) void batch_rhs_kernel(long nd, const int* __restrict__ inst,
                                                        const uint8_t* __restrict__ mode,
                                                        const double* __restrict__ g,
                                                        const double* __restrict__ xsave,
                                                        const double* __restrict__ v0,
                                                        double* __restrict__ b,
                                                        double* __restrict__ dx) {
  const long d = blockIdx.x * 256L + threadIdx.x;
  if (d >= nd) return;
  const int md = mode[inst[d / 3]];
  if (md == 1) {
    b[d] = -g[d];
    dx[d] = xsave[d] - v0[d];
  } else if (md == 2) {
    b[d] = 0.0;
    dx[d] = -g[d];
  } else {
    b[d] = 0.0;
    dx[d] = 0.0;
  }
}

__global__ __launch_bounds__(256) void batch_copy_kernel(long nd, const int* __restrict__ inst,
                                                         const uint8_t* __restrict__ mask,
                                                         const double* __restrict__ src,
                                                         double* __restrict__ dst) {
  const long d = blockIdx.x * 256L + threadIdx.x;
  if (d >= nd) return;
  if (mask[inst[d / 3]]) dst[d] = src[d];
}

static inline unsigned grid_of(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void launch_batch_norm2(hipStream_t s, const BatchArgs& a, const double* v, double* out) {
  if (a.K == 0) return;
  hipLaunchKernelGGL(batch_norm2_kernel, dim3(a.K), dim3(BR_BLOCK), 0, s, a, v, out);
}

void launch_batch_step(hipStream_t s, const BatchArgs& a, long n_nodes, const double* fac,
                       const double* xsave, const double* dx, double* x) {
  const long nd = 3 * n_nodes;
  if (nd == 0) return;
  hipLaunchKernelGGL(batch_step_kernel, dim3(grid_of(nd, 256)), dim3(256), 0, s, nd,
                     a.node_inst, fac, xsave, dx, x);
}

void launch_batch_rhs(hipStream_t s, const BatchArgs& a, long n_nodes, const uint8_t* mode,
                      const double* g, const double* xsave, const double* v0, double* b,
                      double* dx) {
  const long nd = 3 * n_nodes;
  if (nd == 0) return;
  hipLaunchKernelGGL(batch_rhs_kernel, dim3(grid_of(nd, 256)), dim3(256), 0, s, nd,
                     a.node_inst, mode, g, xsave, v0, b, dx);
}

void launch_batch_copy(hipStream_t s, const BatchArgs& a, long n_nodes, const uint8_t* mask,
                       const double* src, double* dst) {
  const long nd = 3 * n_nodes;
  if (nd == 0) return;
  hipLaunchKernelGGL(batch_copy_kernel, dim3(grid_of(nd, 256)), dim3(256), 0, s, nd,
                     a.node_inst, mask, src, dst);
}

}  // namespace pgn

namespace pgn {

__global__ __launch_bounds__(BR_BLOCK) void batch_dot_kernel(BatchArgs A,
                                                             const double* __restrict__ x,
                                                             const double* __restrict__ y,
                                                             double* __restrict__ out) {
  __shared__ double red[BR_BLOCK / 64];
  const int k = blockIdx.x;
  const long d0 = 3 * A.noff[k], d1 = 3 * A.noff[k + 1];
  double s = 0.0;
  for (long d = d0 + threadIdx.x; d < d1; d += BR_BLOCK) s += x[d] * y[d];
  s = block_sum<BR_BLOCK>(s, red);
  if (threadIdx.x == 0) out[k] = s;
}

template <class T>
__global__ __launch_bounds__(256) void batch_add_diag_kernel(long n, const int* __restrict__ inst,
                                                             const long* __restrict__ diag,
                                                             T* __restrict__ bval,
                                                             const double* __restrict__ shift) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= n) return;
  const double sh = shift[inst[i]];
  if (sh == 0.0) return;
  T* B = bval + 9 * diag[i];
  B[0] += (T)sh;
  B[4] += (T)sh;
  B[8] += (T)sh;
}

void launch_batch_dot(hipStream_t s, const BatchArgs& a, const double* x, const double* y,
                      double* out) {
  if (a.K == 0) return;
  hipLaunchKernelGGL(batch_dot_kernel, dim3(a.K), dim3(BR_BLOCK), 0, s, a, x, y, out);
}

void launch_batch_add_diag(hipStream_t s, bool fp32, const BatchArgs& a, long n_nodes,
                           const long* diag, void* bval, const double* shift) {
  if (n_nodes == 0) return;
  const unsigned g = (unsigned)((n_nodes + 255) / 256);
  if (fp32)
    hipLaunchKernelGGL(batch_add_diag_kernel<float>, dim3(g), dim3(256), 0, s, n_nodes,
                       a.node_inst, diag, reinterpret_cast<float*>(bval), shift);
  else
    hipLaunchKernelGGL(batch_add_diag_kernel<double>, dim3(g), dim3(256), 0, s, n_nodes,
                       a.node_inst, diag, reinterpret_cast<double*>(bval), shift);
}

}  // namespace pgn

namespace pgn {

// |A_k|_F^2 over the stored lower entry set of each instance's rows (node
// blocks with row >= col), one workgroup per instance
template <class T>
__global__ __launch_bounds__(BR_BLOCK) void batch_frob_kernel(BatchArgs A,
                                                              const long* __restrict__ rowptr,
                                                              const int* __restrict__ brow,
                                                              const int* __restrict__ bcol,
                                                              const T* __restrict__ bval,
                                                              double* __restrict__ out) {
  __shared__ double red[BR_BLOCK / 64];
  const int k = blockIdx.x;
  const long s0 = rowptr[A.noff[k]], s1 = rowptr[A.noff[k + 1]];
  double acc = 0.0;
  for (long s = s0 + threadIdx.x; s < s1; s += BR_BLOCK)
    if (brow[s] >= bcol[s]) {
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        const double v = bval[s * 9 + q];
        acc += v * v;
      }
    }
  acc = block_sum<BR_BLOCK>(acc, red);
  if (threadIdx.x == 0) out[k] = acc;
}

void launch_batch_frob(hipStream_t s, bool fp32, const BatchArgs& a, const long* rowptr,
                       const int* brow, const int* bcol, const void* bval, double* out) {
  if (a.K == 0) return;
  if (fp32)
    hipLaunchKernelGGL(batch_frob_kernel<float>, dim3(a.K), dim3(BR_BLOCK), 0, s, a, rowptr,
                       brow, bcol, reinterpret_cast<const float*>(bval), out);
  else
    hipLaunchKernelGGL(batch_frob_kernel<double>, dim3(a.K), dim3(BR_BLOCK), 0, s, a, rowptr,
                       brow, bcol, reinterpret_cast<const double*>(bval), out);
}

}  // namespace pgn

namespace pgn {

__global__ __launch_bounds__(256) void batch_axpy_f32_kernel(long nd, const int* __restrict__ inst,
                                                             const uint8_t* __restrict__ mask,
                                                             const float* __restrict__ d,
                                                             double* __restrict__ x) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= nd) return;
  if (mask[inst[i / 3]]) x[i] += (double)d[i];
}

void launch_batch_axpy_f32(hipStream_t s, const BatchArgs& a, long n_nodes, const uint8_t* mask,
                           const float* d, double* x) {
  const long nd = 3 * n_nodes;
  if (nd == 0) return;
  hipLaunchKernelGGL(batch_axpy_f32_kernel, dim3(grid_of(nd, 256)), dim3(256), 0, s, nd,
                     a.node_inst, mask, d, x);
}

}  // namespace pgn

namespace pgn {

constexpr int SR_BS = 256;

__global__ __launch_bounds__(SR_BS) void seg_dot_kernel(SegRedTable t,
                                                        const uint8_t* __restrict__ mask,
                                                        const double* __restrict__ x,
                                                        const double* __restrict__ y) {
  __shared__ double red[SR_BS / 64];
  const int w = blockIdx.x, k = t.inst[w];
  if (mask && !mask[k]) return;
  double v = 0.0;
  const long hi = t.hi[w];
  for (long d = t.lo[w] + threadIdx.x; d < hi; d += SR_BS) v += x[d] * y[d];
  v = block_sum<SR_BS>(v, red);
  if (threadIdx.x == 0) t.part[w] = v;
}

template <class T>
__global__ __launch_bounds__(SR_BS) void seg_frob_kernel(SegRedTable t,
                                                         const uint8_t* __restrict__ mask,
                                                         const int* __restrict__ brow,
                                                         const int* __restrict__ bcol,
                                                         const T* __restrict__ bval) {
  __shared__ double red[SR_BS / 64];
  const int w = blockIdx.x, k = t.inst[w];
  if (mask && !mask[k]) return;
  double v = 0.0;
  const long hi = t.hi[w];
  for (long b = t.lo[w] + threadIdx.x; b < hi; b += SR_BS)
    if (brow[b] >= bcol[b]) {
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        const double e = bval[b * 9 + q];
        v += e * e;
      }
    }
  v = block_sum<SR_BS>(v, red);
  if (threadIdx.x == 0) t.part[w] = v;
}

__global__ __launch_bounds__(SR_BS) void seg_energy_kernel(SegRedTable t,
                                                           const uint8_t* __restrict__ mask,
                                                           const double* __restrict__ elemE,
                                                           const double* __restrict__ nodeE,
                                                           const double* __restrict__ nodeG2) {
  __shared__ double red[2 * SR_BS / 64];
  const int w = blockIdx.x, k = t.inst[w];
  if (mask && !mask[k]) return;
  double el = 0.0, ld = 0.0, gg = 0.0;
  const long hi = t.hi[w];
  if (t.kind[w] == 0) {
    for (long e = t.lo[w] + threadIdx.x; e < hi; e += SR_BS) el += elemE[e];
  } else {
    for (long i = t.lo[w] + threadIdx.x; i < hi; i += SR_BS) {
      ld += nodeE[i];
      gg += nodeG2[i];
    }
  }
  block_sum2<SR_BS>(el, ld, red);
  gg = block_sum<SR_BS>(gg, red);
  if (threadIdx.x == 0) {
    t.part[3L * w] = el;
    t.part[3L * w + 1] = ld;
    t.part[3L * w + 2] = gg;
  }
}

__global__ __launch_bounds__(64) void seg_energy_fin_kernel(int K, SegRedTable t,
                                                            const uint8_t* __restrict__ mask,
                                                            double* __restrict__ out) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  if (k >= K || (mask && !mask[k])) return;
  double el = 0.0, ld = 0.0, gg = 0.0;
  for (int w = t.wg0[k]; w < t.wg0[k + 1]; ++w) {
    el += t.part[3L * w];
    ld += t.part[3L * w + 1];
    gg += t.part[3L * w + 2];
  }
  out[4 * k] = el + ld;
  out[4 * k + 1] = gg;
  out[4 * k + 2] = fabs(el) + fabs(ld);
  out[4 * k + 3] = el;
}

void launch_seg_energy(hipStream_t s, int K, const SegRedTable& t, const uint8_t* mask,
                       const double* elemE, const double* nodeE, const double* nodeG2,
                       double* out) {
  if (K == 0 || t.W == 0) return;
  hipLaunchKernelGGL(seg_energy_kernel, dim3(t.W), dim3(SR_BS), 0, s, t, mask, elemE, nodeE,
                     nodeG2);
  hipLaunchKernelGGL(seg_energy_fin_kernel, dim3((K + 63) / 64), dim3(64), 0, s, K, t, mask,
                     out);
}

// one thread per instance: its partials in order
__global__ __launch_bounds__(64) void seg_red_fin_kernel(int K, SegRedTable t,
                                                         const uint8_t* __restrict__ mask,
                                                         double* __restrict__ out) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  if (k >= K || (mask && !mask[k])) return;
  double v = 0.0;
  for (int w = t.wg0[k]; w < t.wg0[k + 1]; ++w) v += t.part[w];
  out[k] = v;
}

void launch_seg_dot(hipStream_t s, int K, const SegRedTable& t, const uint8_t* mask,
                    const double* x, const double* y, double* out) {
  if (K == 0 || t.W == 0) return;
  hipLaunchKernelGGL(seg_dot_kernel, dim3(t.W), dim3(SR_BS), 0, s, t, mask, x, y);
  hipLaunchKernelGGL(seg_red_fin_kernel, dim3((K + 63) / 64), dim3(64), 0, s, K, t, mask, out);
}

void launch_seg_frob(hipStream_t s, bool fp32, int K, const SegRedTable& t, const uint8_t* mask,
                     const int* brow, const int* bcol, const void* bval, double* out) {
  if (K == 0 || t.W == 0) return;
  if (fp32)
    hipLaunchKernelGGL(seg_frob_kernel<float>, dim3(t.W), dim3(SR_BS), 0, s, t, mask, brow, bcol,
                       reinterpret_cast<const float*>(bval));
  else
    hipLaunchKernelGGL(seg_frob_kernel<double>, dim3(t.W), dim3(SR_BS), 0, s, t, mask, brow,
                       bcol, reinterpret_cast<const double*>(bval));
  hipLaunchKernelGGL(seg_red_fin_kernel, dim3((K + 63) / 64), dim3(64), 0, s, K, t, mask, out);
}

}  // namespace pgn
