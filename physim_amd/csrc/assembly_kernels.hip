// assembly_kernels.hip -- atomic-free gathers of element gradients and
// Hessians into the global gradient and the BSR3 Hessian.
//
// Reference: Simulable::AddGradient / AddHessian (Simulable.cpp:357-701) scatter
// element blocks serially through cached pointers (DerivativeAssembler.cpp:
// 171-301, BasicTypes.cpp:18-164).  Here every output value is owned by one
// lane, which sums its contributions in element order from a slot map built
// once at model setup: deterministic, no atomics, no serial pass.
#include "kernels.hpp"
#include "reduce.hpp"

namespace pgn {

constexpr int NODE_BLOCK = 256;
constexpr int FINAL_BLOCK = 1024;

// Workgroups of the gathers: rounded up to whole rounds of the 8 XCDs, which
// take workgroups round-robin by index.  xcd_block() maps workgroup b to
// logical block (b % 8) * (G / 8) + b / 8, so each XCD walks one contiguous
// eighth of the nodes (blocks) and the element records they share stay in
// that XCD's L2 instead of being fetched by all eight.
static inline int round_up8(long g) { return g >= 8 ? (int)((g + 7) / 8 * 8) : (int)g; }
int node_gather_blocks(int n) { return round_up8((n + NODE_BLOCK - 1) / NODE_BLOCK); }

__device__ __forceinline__ long xcd_block() {
  const long G = gridDim.x, b = blockIdx.x;
  return (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
}

// Gradient = sum of element gradients (element order) - M_i g + f, fixed DoFs
// zeroed (FixVectorDoFs, Simulable.cpp:735-744).  Node energy terms:
// x.(-M g) (MassElement_Lumped.cpp:56-68) and x.f (EnergyElement_Force.cpp:43).
template <class T>
__global__ __launch_bounds__(NODE_BLOCK) void node_gather_kernel(
    NodeArgs A, const T* __restrict__ eg, double* __restrict__ g,
    double* __restrict__ part) {
  __shared__ double red[2 * NODE_BLOCK / 64];
  const long i = A.i0 + xcd_block() * NODE_BLOCK + threadIdx.x;
  double eN = 0.0, gg = 0.0;
  if (i < A.i0 + A.n) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    const int p1 = A.inc_ptr[i + 1];
    int p = A.inc_ptr[i];
    // eight incidences per round: their indices, then their gradient rows,
    // are all in flight before the sums (in incidence order) need them
    constexpr int U = 8;
    for (; p + U <= p1; p += U) {
      int c[U];
#pragma unroll
      for (int j = 0; j < U; ++j) c[j] = A.inc[p + j];
      T v[U][3];
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const T* ge = eg + (long)(c[j] >> 2) * 12 + 3 * (c[j] & 3);
        v[j][0] = ge[0];
        v[j][1] = ge[1];
        v[j][2] = ge[2];
      }
#pragma unroll
      for (int j = 0; j < U; ++j) {
        s0 += (double)v[j][0];
        s1 += (double)v[j][1];
        s2 += (double)v[j][2];
      }
    }
    for (; p < p1; ++p) {
      const int c = A.inc[p];
      const T* ge = eg + (long)(c >> 2) * 12 + 3 * (c & 3);
      s0 += (double)ge[0];
      s1 += (double)ge[1];
      s2 += (double)ge[2];
    }
    const double M = A.mass[i];
    double gv[3] = {A.grav[0], A.grav[1], A.grav[2]};
    double fa = A.f_alpha;
    if (A.node_inst) {
      const int k = A.node_inst[i];
      gv[0] = A.grav_inst[3 * k];
      gv[1] = A.grav_inst[3 * k + 1];
      gv[2] = A.grav_inst[3 * k + 2];
      fa = A.falpha_inst[k];
    }
    double gr[3] = {s0 - M * gv[0], s1 - M * gv[1], s2 - M * gv[2]};
    const double* xi = A.x + 3L * i;
#pragma unroll
    for (int k = 0; k < 3; ++k) eN += xi[k] * (-M * gv[k]);
    if (A.f_end) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double f0 = A.f_ini[3L * i + k];
        const double f = f0 + (A.f_end[3L * i + k] - f0) * fa;
        gr[k] += f;
        eN += xi[k] * f;
      }
    }
    if (A.xhat) {
      // OptimProblem_BasicDynamic.cpp:68-107: with a = (x - xhat) / dt^2,
      // E += dt^2/2 a.M a and g += M a, before the fixed DoFs are zeroed
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double a = (xi[k] - A.xhat[3L * i + k]) * A.inv_dt2;
        gr[k] += M * a;
        eN += 0.5 / A.inv_dt2 * a * (M * a);
      }
    }
    const unsigned fm = A.fixmask[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (fm & (1u << k)) gr[k] = 0.0;
      g[3L * i + k] = gr[k];
      gg += gr[k] * gr[k];
    }
    if (A.nodeE) {
      A.nodeE[i] = eN;
      A.nodeG2[i] = gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2];
    }
  }
  block_sum2<NODE_BLOCK>(eN, gg, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = eN;
    part[2 * blockIdx.x + 1] = gg;
  }
}

void launch_node_gather(hipStream_t s, bool fp32_eg, const NodeArgs& a,
                        const void* eg, double* g, double* node_partials) {
  const int grid = node_gather_blocks(a.n);
  if (grid == 0) return;
  if (fp32_eg)
    hipLaunchKernelGGL(node_gather_kernel<float>, dim3(grid), dim3(NODE_BLOCK), 0, s,
                       a, reinterpret_cast<const float*>(eg), g, node_partials);
  else
    hipLaunchKernelGGL(node_gather_kernel<double>, dim3(grid), dim3(NODE_BLOCK), 0, s,
                       a, reinterpret_cast<const double*>(eg), g, node_partials);
}

__global__ __launch_bounds__(FINAL_BLOCK) void energy_final_kernel(
    const double* __restrict__ ep, int ne, const double* __restrict__ np, int nn,
    double* __restrict__ out2) {
  __shared__ double red[2 * FINAL_BLOCK / 64];
  double e = 0.0, en = 0.0, gg = 0.0;
  for (int i = threadIdx.x; i < ne; i += FINAL_BLOCK) e += ep[i];
  for (int i = threadIdx.x; i < nn; i += FINAL_BLOCK) {
    en += np[2 * i];
    gg += np[2 * i + 1];
  }
  block_sum2<FINAL_BLOCK>(e, en, red);
  const double sg = block_sum<FINAL_BLOCK>(gg, red);
  if (threadIdx.x == 0) {
    out2[0] = e + en;                // elastic + load energy
    out2[1] = sg;                    // |g|^2
    out2[2] = fabs(e) + fabs(en);    // scale of the rounding of out2[0]
  }
}

void launch_energy_final(hipStream_t s, const double* elem_partials,
                         int n_elem_partials, const double* node_partials,
                         int n_node_partials, double* out2) {
  hipLaunchKernelGGL(energy_final_kernel, dim3(1), dim3(FINAL_BLOCK), 0, s,
                     elem_partials, n_elem_partials, node_partials,
                     n_node_partials, out2);
}

// ---- Hessian gather ----------------------------------------------------------
// Block (i, j) of the BSR3 matrix sums, in element order, the packed element
// blocks of every element holding both nodes: local pair (a, b) with a > b is
// stored as is, a < b is the transpose of the stored (b, a), a == b is a
// symmetric diagonal block.  Then FixMatrixDoFs (Simulable.cpp:746-760): rows
// and columns of fixed DoFs become identity.
template <class T>
__global__ __launch_bounds__(256) void bsr_gather_kernel(BsrArgs A,
                                                         const T* __restrict__ eH,
                                                         T* __restrict__ bval) {
  const long s = A.s0 + xcd_block() * 256L + threadIdx.x;
  if (s >= A.s0 + A.nb) return;
  T acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) acc[q] = T(0);
  const long c1 = A.cptr[s + 1];
  long p = A.cptr[s];
  // GU contributions at a time: their codes, then all their value loads are
  // issued before any is added (the adds stay in element order)
  constexpr int GU = 4;
  for (; p + GU <= c1; p += GU) {
    uint32_t cc[GU];
#pragma unroll
    for (int u = 0; u < GU; ++u) cc[u] = A.code[p + u];
    T v[GU][9];
#pragma unroll
    for (int u = 0; u < GU; ++u) {
      const long e = cc[u] >> 4;
      const int la = (cc[u] >> 2) & 3, lb = cc[u] & 3;
      const T* __restrict__ h = eH + e * TET_H_PACK;
      if (la == lb) {
        const T* d = h + 54 + 6 * la;
#pragma unroll
        for (int q = 0; q < 6; ++q) v[u][q] = d[q];
      } else {
        const int hi = la > lb ? la : lb, lo = la > lb ? lb : la;
        const T* o = h + 9 * (hi * (hi - 1) / 2 + lo);
#pragma unroll
        for (int q = 0; q < 9; ++q) v[u][q] = o[q];
      }
    }
#pragma unroll
    for (int u = 0; u < GU; ++u) {
      const int la = (cc[u] >> 2) & 3, lb = cc[u] & 3;
      if (la == lb) {
        const T d0 = v[u][0], d1 = v[u][1], d2 = v[u][2], d3 = v[u][3], d4 = v[u][4],
                d5 = v[u][5];
        acc[0] += d0; acc[1] += d1; acc[2] += d3;
        acc[3] += d1; acc[4] += d2; acc[5] += d4;
        acc[6] += d3; acc[7] += d4; acc[8] += d5;
      } else if (la > lb) {
#pragma unroll
        for (int q = 0; q < 9; ++q) acc[q] += v[u][q];
      } else {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int k = 0; k < 3; ++k) acc[3 * i + k] += v[u][3 * k + i];
      }
    }
  }
  for (; p < c1; ++p) {
    const uint32_t c = A.code[p];
    const long e = c >> 4;
    const int la = (c >> 2) & 3, lb = c & 3;
    const T* __restrict__ h = eH + e * TET_H_PACK;
    if (la == lb) {
      const T* d = h + 54 + 6 * la;
      const T d0 = d[0], d1 = d[1], d2 = d[2], d3 = d[3], d4 = d[4], d5 = d[5];
      acc[0] += d0; acc[1] += d1; acc[2] += d3;
      acc[3] += d1; acc[4] += d2; acc[5] += d4;
      acc[6] += d3; acc[7] += d4; acc[8] += d5;
    } else if (la > lb) {
      const T* o = h + 9 * (la * (la - 1) / 2 + lb);
#pragma unroll
      for (int q = 0; q < 9; ++q) acc[q] += o[q];
    } else {
      const T* o = h + 9 * (lb * (lb - 1) / 2 + la);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[3 * i + k] += o[3 * k + i];
    }
  }
  const int bi = A.brow[s], bj = A.bcol[s];
  if (A.mass && bi == bj) {
    // OptimProblem_BasicDynamic.cpp:115-119: H += M / dt^2
    const T md = (T)(A.mass[bi] * A.inv_dt2);
    acc[0] += md;
    acc[4] += md;
    acc[8] += md;
  }
  const unsigned fi = A.fixmask[bi], fj = A.fixmask[bj];
  if (fi | fj) {
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q)
        if (((fi >> p) & 1u) || ((fj >> q) & 1u))
          acc[3 * p + q] = (bi == bj && p == q) ? T(1) : T(0);
  }
  T* o = bval + s * 9;
#pragma unroll
  for (int q = 0; q < 9; ++q) o[q] = acc[q];
}

void launch_bsr_gather(hipStream_t s, bool fp32, const BsrArgs& a,
                       const void* eH, void* bval) {
  const long grid = round_up8((a.nb + 255) / 256);
  if (grid == 0) return;
  if (fp32)
    hipLaunchKernelGGL(bsr_gather_kernel<float>, dim3(grid), dim3(256), 0, s, a,
                       reinterpret_cast<const float*>(eH), reinterpret_cast<float*>(bval));
  else
    hipLaunchKernelGGL(bsr_gather_kernel<double>, dim3(grid), dim3(256), 0, s, a,
                       reinterpret_cast<const double*>(eH), reinterpret_cast<double*>(bval));
}

// Jacobi: Eigen's DiagonalPreconditioner (1/a_ii, or 1 where a_ii == 0).
// Block Jacobi: inverse of each 3x3 diagonal node block (Jacobi where the
// block is singular), stored as its upper triangle {m00, m01, m02, m11, m12,
// m22}: diagonal blocks are assembled exactly symmetric, and the cofactor
// of (i, j) then takes the same two products as that of (j, i), so the
// inverse is exactly symmetric too.
template <class T>
__global__ __launch_bounds__(256) void precond_kernel(int n, const long* __restrict__ diag,
                                                     const T* __restrict__ bval,
                                                     T* __restrict__ dinv,
                                                     T* __restrict__ binv, long i0) {
  const long i = i0 + blockIdx.x * 256L + threadIdx.x;
  if (i >= i0 + n) return;
  const T* B = bval + diag[i] * 9;
  T a[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) a[q] = B[q];
#pragma unroll
  for (int p = 0; p < 3; ++p) dinv[3L * i + p] = (a[4 * p] != T(0)) ? T(1) / a[4 * p] : T(1);
  const double c00 = (double)a[4] * a[8] - (double)a[5] * a[7];
  const double c01 = (double)a[5] * a[6] - (double)a[3] * a[8];
  const double c02 = (double)a[3] * a[7] - (double)a[4] * a[6];
  const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  T* o = binv + 6L * i;
  if (det != 0.0 && isfinite(det)) {
    const double id = 1.0 / det;
    o[0] = (T)(c00 * id);
    o[1] = (T)(((double)a[2] * a[7] - (double)a[1] * a[8]) * id);
    o[2] = (T)(((double)a[1] * a[5] - (double)a[2] * a[4]) * id);
    o[3] = (T)(((double)a[0] * a[8] - (double)a[2] * a[6]) * id);
    o[4] = (T)(((double)a[2] * a[3] - (double)a[0] * a[5]) * id);
    o[5] = (T)(((double)a[0] * a[4] - (double)a[1] * a[3]) * id);
  } else {
    o[0] = dinv[3L * i];
    o[1] = o[2] = o[4] = T(0);
    o[3] = dinv[3L * i + 1];
    o[5] = dinv[3L * i + 2];
  }
}

void launch_precond(hipStream_t s, bool fp32, int n, const long* diag,
                    const void* bval, void* dinv, void* binv, long i0) {
  const int grid = (n + 255) / 256;
  if (grid <= 0) return;
  if (fp32)
    hipLaunchKernelGGL(precond_kernel<float>, dim3(grid), dim3(256), 0, s, n, diag,
                       reinterpret_cast<const float*>(bval), reinterpret_cast<float*>(dinv),
                       reinterpret_cast<float*>(binv), i0);
  else
    hipLaunchKernelGGL(precond_kernel<double>, dim3(grid), dim3(256), 0, s, n, diag,
                       reinterpret_cast<const double*>(bval), reinterpret_cast<double*>(dinv),
                       reinterpret_cast<double*>(binv), i0);
}

template <class T>
__global__ __launch_bounds__(256) void add_diag_kernel(int n, const long* __restrict__ diag,
                                                      T* __restrict__ bval, T shift) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  T* B = bval + diag[i] * 9;
  B[0] += shift;
  B[4] += shift;
  B[8] += shift;
}

void launch_add_diag(hipStream_t s, bool fp32, int n, const long* diag,
                     void* bval, double shift) {
  const int grid = (n + 255) / 256;
  if (grid == 0) return;
  if (fp32)
    hipLaunchKernelGGL(add_diag_kernel<float>, dim3(grid), dim3(256), 0, s, n, diag,
                       reinterpret_cast<float*>(bval), (float)shift);
  else
    hipLaunchKernelGGL(add_diag_kernel<double>, dim3(grid), dim3(256), 0, s, n, diag,
                       reinterpret_cast<double*>(bval), shift);
}

// |A|_F^2 over the entry set AMatrixSd stores: node blocks with row >= col,
// diagonal blocks whole (LinearSolver.cpp:57-66 uses mA.norm()).
constexpr int FROB_GRID = 512;
int frob_blocks() { return FROB_GRID; }

template <class T>
__global__ __launch_bounds__(256) void frob_kernel(long nb, const int* __restrict__ brow,
                                                  const int* __restrict__ bcol,
                                                  const T* __restrict__ bval,
                                                  double* __restrict__ part,
                                                  double* __restrict__ out,
                                                  unsigned* ticket) {
  __shared__ double red[4];
  __shared__ bool last;
  double acc = 0.0;
  for (long s = blockIdx.x * 256L + threadIdx.x; s < nb; s += (long)gridDim.x * 256) {
    if (brow[s] >= bcol[s]) {
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        const double v = bval[s * 9 + q];
        acc += v * v;
      }
    }
  }
  const double bs = block_sum<256>(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
  if (last_block(ticket, &last)) {
    const double t = block_sum_array<256>(part, gridDim.x, 1, red);
    if (threadIdx.x == 0) out[0] = t;
  }
}

void launch_frob_lower(hipStream_t s, bool fp32, long nb, const int* brow,
                       const int* bcol, const void* bval, double* partials,
                       double* out) {
  // partials holds PCG_MAXP + 4 doubles; this kernel's ticket sits past
  // the ticket of the residual / dot kernels (zeroed at allocation and reset
  // by the last workgroup).
  unsigned* ticket = reinterpret_cast<unsigned*>(partials + PCG_MAXP + 2);
  if (fp32)
    hipLaunchKernelGGL(frob_kernel<float>, dim3(FROB_GRID), dim3(256), 0, s, nb, brow,
                       bcol, reinterpret_cast<const float*>(bval), partials, out, ticket);
  else
    hipLaunchKernelGGL(frob_kernel<double>, dim3(FROB_GRID), dim3(256), 0, s, nb, brow,
                       bcol, reinterpret_cast<const double*>(bval), partials, out, ticket);
}

}  // namespace pgn
