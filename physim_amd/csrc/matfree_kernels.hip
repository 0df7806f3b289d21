// matfree_kernels.hip -- the Hessian product without the assembled matrix:
// q = H p as the sum over tets of their element products (the product the
// PCG iteration forms, LinearSolver_EigenCG.cpp:89 through Eigen's CG).
//
// Two launches, deterministic (no atomics):
//   element: y_e = H_e p_e from the stored packed element Hessian (78
//            values: 6 off-diagonal 3x3 node blocks a > b, then 4 diagonal
//            blocks as lower triangles), p with its fixed DoFs zeroed (the
//            fixed columns of the projected Hessian are zero); a workgroup's
//            packed Hessians are staged through LDS so that the loads are
//            contiguous 16-byte vectors instead of 64 lanes 312 B apart;
//   node:    q_i = sum of y_e over the node's incidences in element order,
//            fixed rows q = p (identity), plus the regularization shift
//            (A += shift I), and the p.q partials of the PCG.
// Bytes per product, FP32: 312 (H) + 16 (ids) + 48 (y) per tet, 12 per
// incidence (y read back) + 4 per incidence (its code) + 3 x 12 per node (p
// twice, q): about 3x the assembled SELL-64 product on a tet mesh.
#include <type_traits>

#include "kernels.hpp"
#include "reduce.hpp"

namespace pgn {

constexpr int MF_BS = 128;  // tets per element workgroup (39 KB of FP32 staging)

template <class T>
__global__ __launch_bounds__(MF_BS) void mf_elem_kernel(long m, const int4* __restrict__ tets,
                                                        const T* __restrict__ eH,
                                                        const T* __restrict__ p,
                                                        const uint8_t* __restrict__ fix,
                                                        T* __restrict__ ye) {
  __shared__ __attribute__((aligned(16))) T hs[MF_BS * TET_H_PACK];
  const long e0 = (long)blockIdx.x * MF_BS;
  const int ne = (int)(m - e0 < MF_BS ? m - e0 : MF_BS);
  {
    // the workgroup's packed Hessians: one contiguous stretch
    using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    constexpr int W = 16 / sizeof(T);
    const int nv = ne * TET_H_PACK, nw = nv / W;
    const V* src = reinterpret_cast<const V*>(eH + e0 * TET_H_PACK);
    V* dst = reinterpret_cast<V*>(hs);
    for (int q = threadIdx.x; q < nw; q += MF_BS) dst[q] = src[q];
    for (int q = W * nw + threadIdx.x; q < nv; q += MF_BS) hs[q] = eH[e0 * TET_H_PACK + q];
  }
  __syncthreads();
  const long e = e0 + threadIdx.x;
  if (e >= m) return;
  const int4 t = tets[e];
  const int nd[4] = {t.x, t.y, t.z, t.w};
  T pv[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const uint8_t f = fix[nd[a]];
#pragma unroll
    for (int c = 0; c < 3; ++c) pv[a][c] = (f >> c) & 1 ? T(0) : p[3L * nd[a] + c];
  }
  const T* h = hs + threadIdx.x * TET_H_PACK;
  T y[4][3];
  // diagonal blocks (lower triangles d0 d1 d2 d3 d4 d5 = 00 10 11 20 21 22)
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const T* d = h + 54 + 6 * a;
    y[a][0] = d[0] * pv[a][0] + d[1] * pv[a][1] + d[3] * pv[a][2];
    y[a][1] = d[1] * pv[a][0] + d[2] * pv[a][1] + d[4] * pv[a][2];
    y[a][2] = d[3] * pv[a][0] + d[4] * pv[a][1] + d[5] * pv[a][2];
  }
  // off-diagonal blocks B_ab (a > b, row-major): y_a += B p_b, y_b += B^T p_a
#pragma unroll
  for (int a = 1; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < a; ++b) {
      const T* B = h + 9 * (a * (a - 1) / 2 + b);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        y[a][i] += B[3 * i] * pv[b][0] + B[3 * i + 1] * pv[b][1] + B[3 * i + 2] * pv[b][2];
        y[b][i] += B[i] * pv[a][0] + B[3 + i] * pv[a][1] + B[6 + i] * pv[a][2];
      }
    }
  T* o = ye + e * 12;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) o[3 * a + c] = y[a][c];
}

// Rows of node i gathered from the element products; the workgroup's rows
// are the PCG SpMV's (same grid, same XCD-blocked partition) so that its
// p.q partial lands where the PCG update re-sums it.
template <class T>
__global__ __launch_bounds__(256) void mf_node_kernel(int n, const int* __restrict__ inc_ptr,
                                                      const int* __restrict__ inc,
                                                      const T* __restrict__ ye,
                                                      const uint8_t* __restrict__ fix,
                                                      T shift, const T* __restrict__ p,
                                                      T* __restrict__ q, double* part) {
  __shared__ double red[256 / 64];
  const int G = gridDim.x, b = blockIdx.x;
  const int lb = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
  const long chunk = ((long)n + G - 1) / G;
  const long lo = (long)lb * chunk, hi = lo + chunk < n ? lo + chunk : n;
  double pq = 0.0;
  for (long i = lo + threadIdx.x; i < hi; i += 256) {
    T y0 = 0, y1 = 0, y2 = 0;
    const int k0 = inc_ptr[i], k1 = inc_ptr[i + 1];
    // eight incidences' loads in flight, summed in element order
    for (int k = k0; k < k1; k += 8) {
      int c[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) c[u] = k + u < k1 ? inc[k + u] : -1;
      T v[8][3];
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int d = 0; d < 3; ++d) v[u][d] = c[u] >= 0 ? ye[3L * c[u] + d] : T(0);
#pragma unroll
      for (int u = 0; u < 8; ++u) {  // element * 4 + local node -> 3 values
        y0 += v[u][0];
        y1 += v[u][1];
        y2 += v[u][2];
      }
    }
    const uint8_t f = fix[i];
    const T p0 = p[3 * i], p1 = p[3 * i + 1], p2 = p[3 * i + 2];
    y0 = (f & 1) ? p0 + shift * p0 : y0 + shift * p0;
    y1 = (f & 2) ? p1 + shift * p1 : y1 + shift * p1;
    y2 = (f & 4) ? p2 + shift * p2 : y2 + shift * p2;
    q[3 * i] = y0;
    q[3 * i + 1] = y1;
    q[3 * i + 2] = y2;
    pq += (double)p0 * (double)y0 + (double)p1 * (double)y1 + (double)p2 * (double)y2;
  }
  const double s = block_sum<256>(pq, red);
  if (threadIdx.x == 0 && part) part[blockIdx.x] = s;
}

void launch_matfree_product(hipStream_t s, bool fp32, const MatFreeArgs& a, const void* p,
                            void* q, double* part, int grid) {
  const int ge = (int)((a.m + MF_BS - 1) / MF_BS);
  if (ge > 0) {
    if (fp32)
      hipLaunchKernelGGL(mf_elem_kernel<float>, dim3(ge), dim3(MF_BS), 0, s, a.m, a.tets,
                         reinterpret_cast<const float*>(a.eH),
                         reinterpret_cast<const float*>(p), a.fixmask,
                         reinterpret_cast<float*>(a.ye));
    else
      hipLaunchKernelGGL(mf_elem_kernel<double>, dim3(ge), dim3(MF_BS), 0, s, a.m, a.tets,
                         reinterpret_cast<const double*>(a.eH),
                         reinterpret_cast<const double*>(p), a.fixmask,
                         reinterpret_cast<double*>(a.ye));
  }
  if (grid < 1) grid = 1;
  if (fp32)
    hipLaunchKernelGGL(mf_node_kernel<float>, dim3(grid), dim3(256), 0, s, a.n, a.inc_ptr,
                       a.inc, reinterpret_cast<const float*>(a.ye), a.fixmask,
                       (float)a.shift, reinterpret_cast<const float*>(p),
                       reinterpret_cast<float*>(q), part);
  else
    hipLaunchKernelGGL(mf_node_kernel<double>, dim3(grid), dim3(256), 0, s, a.n, a.inc_ptr,
                       a.inc, reinterpret_cast<const double*>(a.ye), a.fixmask, a.shift,
                       reinterpret_cast<const double*>(p), reinterpret_cast<double*>(q), part);
}

}  // namespace pgn
