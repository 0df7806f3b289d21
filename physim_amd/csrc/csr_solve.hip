// csr_solve.hip -- ILinearSolver::Solve(A, b, x) on a matrix the caller
// assembled (LinearSolver.h:38-52), without a device model:
// LinearSolver_EigenCG::SolveInternal (LinearSolver_EigenCG.cpp:51-125) in
// FP64 over a scalar CSR matrix.
//
// Per regularization try: x restarts from the caller's x, Jacobi PCG with
// Eigen's ConjugateGradient semantics (DiagonalPreconditioner: 1 / a_ii, or 1
// where a_ii is 0; stop when |r|^2 < max(tol^2 |b|^2, DBL_MIN); the count
// of completed iterations is Eigen's), then the indefiniteness test b.x < 0;
// try i > 0 first adds regT 10^(i-1) to every diagonal entry, cumulatively,
// regT = 1e-6 |A|_F over the stored entries (LinearSolver.cpp:57-66), and
// the shifts stay in the caller's matrix as in the reference (mA is
// modified in place, LinearSolver_EigenCG.cpp:73).
//
// Device layout: the full symmetric matrix as row-major CSR (i64 row
// pointers, i32 columns, FP64 values).  Rows are cut into G <= 1024
// contiguous ranges, one per workgroup and the same in every launch, blocked
// per XCD; L lanes per row (a power of two from the mean row length) walk a
// row's entries in order and reduce by shuffles.  Dot products are
// fence-free: each workgroup writes one partial, every consumer workgroup of
// the next launch re-sums all G partials in the same fixed order, so every
// run of the same inputs repeats bit for bit.  One PCG iteration is three
// launches (SpMV + p.q; alpha, x, r, z and the r.r, r.z partials; the
// convergence test and p = z + beta p); the host queues them in chunks and
// reads the state between chunks.
//
// Two entry points share the kernels: pgn_csr_linear_solve takes host arrays
// and allocates its device copies and work space; pgn_csr_linear_solve_noalloc
// takes device arrays and keeps its work space in a static device array
// (`scratch`) of a fixed maximum size, allocating nothing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "pgn_internal.hpp"
#include "reduce.hpp"

namespace pgn {
int api_guard_set(const char* msg, int code);
}

using namespace pgn;

namespace {

constexpr int CSR_BS = 256;     // threads per workgroup
constexpr int CSR_GMAX = 1024;  // workgroups per launch = partials per reduction
constexpr int CSR_MAX_DEVICES = 64;

struct CsrState {
  double abs_new[2];  // r.z by iteration parity (Eigen's absNew)
  double rr;          // |r|^2 at the last convergence test
  double bb;          // |b|^2
  double thresh;      // max(tol^2 |b|^2, DBL_MIN)
  double normA2;      // sum of squares of the stored entries
  double bx;          // b.x after a try
  int done;           // 1 once the try has converged (or b = 0)
  int iters;          // Eigen's iteration count of the try
  int bad;            // malformed rowptr / column index
  int nodiag;         // a row without a stored diagonal entry
};

// Work space of one solve over n rows, carved from one buffer (256-byte
// aligned pieces): 6 FP64 vectors, the diagonal positions, 3 x CSR_GMAX
// partials and the state.
struct CsrWork {
  double *r, *z, *p, *q, *xprev, *dinv;
  long* dpos;
  double *pa, *pb, *pc;
  CsrState* st;
};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

size_t csr_work_bytes(long n) {
  return 6 * align256(8 * (size_t)n) + align256(8 * (size_t)n) + 3 * align256(8 * CSR_GMAX) +
         align256(sizeof(CsrState));
}

CsrWork csr_carve(char* base, long n) {
  CsrWork w;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* p = base + off;
    off += align256(bytes);
    return p;
  };
  w.r = (double*)take(8 * n);
  w.z = (double*)take(8 * n);
  w.p = (double*)take(8 * n);
  w.q = (double*)take(8 * n);
  w.xprev = (double*)take(8 * n);
  w.dinv = (double*)take(8 * n);
  w.dpos = (long*)take(8 * n);
  w.pa = (double*)take(8 * CSR_GMAX);
  w.pb = (double*)take(8 * CSR_GMAX);
  w.pc = (double*)take(8 * CSR_GMAX);
  w.st = (CsrState*)take(sizeof(CsrState));
  return w;
}

// The allocation-free entry point's work space: rows up to
// PGN_CSR_NOALLOC_MAX_ROWS (14.7 MB on every device the library runs on,
// held until the process ends).  Host code never passes `scratch` itself to
// a kernel: its address on each device comes from hipGetSymbolAddress.
constexpr long CSR_SCRATCH_BYTES =
    6 * ((8L * PGN_CSR_NOALLOC_MAX_ROWS + 255) & ~255L) +
    ((8L * PGN_CSR_NOALLOC_MAX_ROWS + 255) & ~255L) + 3 * ((8L * CSR_GMAX + 255) & ~255L) +
    ((long(sizeof(CsrState)) + 255) & ~255L);
static __device__ __align__(256) char scratch[CSR_SCRATCH_BYTES];

std::mutex g_scratch_mu;
void* g_scratch_ptr[CSR_MAX_DEVICES] = {};

// workgroup -> contiguous row range, consecutive ranges on one XCD
__device__ __forceinline__ int csr_chunk(int b, int G) {
  return (G & 7) == 0 ? (b & 7) * (G >> 3) + (b >> 3) : b;
}

__device__ __forceinline__ void csr_range(int n, int G, long& r0, long& r1) {
  const long rpb = ((long)n + G - 1) / G;
  r0 = (long)csr_chunk(blockIdx.x, G) * rpb;
  r1 = min((long)n, r0 + rpb);
}

// fixed-order re-sum of G partials, broadcast to the whole workgroup
__device__ __forceinline__ double csr_resum(const double* part, int G, double* lds) {
  const double s = block_sum_array<CSR_BS>(part, G, 1, lds);
  if (threadIdx.x == 0) lds[CSR_BS / 64] = s;
  __syncthreads();
  const double r = lds[CSR_BS / 64];
  __syncthreads();
  return r;
}

__global__ void csr_state_init_kernel(CsrState* st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->abs_new[0] = st->abs_new[1] = 0.0;
    st->rr = st->bb = st->thresh = st->normA2 = st->bx = 0.0;
    st->done = st->iters = st->bad = st->nodiag = 0;
  }
}

// Input check, diagonal positions and the partials of |A|_F^2 over the
// stored entries.
__global__ __launch_bounds__(CSR_BS) void csr_setup_kernel(int n, long nnz, int G,
                                                           const long* rp, const int* ci,
                                                           const double* v, long* dpos,
                                                           double* part, CsrState* st) {
  __shared__ double lds[CSR_BS / 64 + 1];
  long r0, r1;
  csr_range(n, G, r0, r1);
  double acc = 0.0;
  bool bad = false, nodiag = false;
  for (long i = r0 + threadIdx.x; i < r1; i += CSR_BS) {
    const long k0 = rp[i], k1 = rp[i + 1];
    long d = -1;
    if (k0 < 0 || k1 < k0 || k1 > nnz || (i == 0 && k0 != 0)) {
      bad = true;
    } else {
      for (long k = k0; k < k1; ++k) {
        const int c = ci[k];
        if (c < 0 || c >= n) {
          bad = true;
          break;
        }
        acc += v[k] * v[k];
        if (c == i && d < 0) d = k;
      }
    }
    if (d < 0) nodiag = true;
    dpos[i] = d;
  }
  if (bad) st->bad = 1;
  if (nodiag) st->nodiag = 1;
  const double tot = block_sum<CSR_BS>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// one workgroup: out = sum of G partials in the fixed order
__global__ __launch_bounds__(CSR_BS) void csr_sum_kernel(const double* part, int G,
                                                         double* out) {
  __shared__ double lds[CSR_BS / 64 + 1];
  const double s = block_sum_array<CSR_BS>(part, G, 1, lds);
  if (threadIdx.x == 0) *out = s;
}

__global__ __launch_bounds__(CSR_BS) void csr_copy_kernel(int n, int G, const double* src,
                                                          double* dst) {
  long r0, r1;
  csr_range(n, G, r0, r1);
  for (long i = r0 + threadIdx.x; i < r1; i += CSR_BS) dst[i] = src[i];
}

// regularization: a_ii += s on every stored diagonal entry
__global__ __launch_bounds__(CSR_BS) void csr_shift_kernel(int n, int G, const long* dpos,
                                                           double* v, double s) {
  long r0, r1;
  csr_range(n, G, r0, r1);
  for (long i = r0 + threadIdx.x; i < r1; i += CSR_BS) v[dpos[i]] += s;
}

// start of a try: x from the caller's x, Jacobi inverse diagonal, state reset
__global__ __launch_bounds__(CSR_BS) void csr_restore_kernel(int n, int G, const double* xprev,
                                                             double* x, const long* dpos,
                                                             const double* v, double* dinv,
                                                             CsrState* st) {
  long r0, r1;
  csr_range(n, G, r0, r1);
  for (long i = r0 + threadIdx.x; i < r1; i += CSR_BS) {
    x[i] = xprev[i];
    const double d = v[dpos[i]];
    dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->done = 0;
    st->iters = 0;
    st->rr = 0.0;
  }
}

// one load of the done flag per workgroup, broadcast through LDS, so every
// wave of the workgroup takes the same branch
__device__ __forceinline__ bool csr_done(const CsrState* st, int* flag) {
  if (threadIdx.x == 0) *flag = st->done;
  __syncthreads();
  return *flag != 0;
}

// q = A u over the block's rows, L lanes per row, entries in stored order;
// returns this thread's share of u.q (lane 0 of each row group)
template <int L>
__device__ __forceinline__ double csr_rows_spmv(long r0, long r1, const long* rp, const int* ci,
                                                const double* v, const double* u, double* q) {
  const int grp = threadIdx.x / L, lane = threadIdx.x % L;
  double acc = 0.0;
  for (long base = r0; base < r1; base += CSR_BS / L) {
    const long i = base + grp;
    const bool ok = i < r1;
    double s = 0.0;
    if (ok)
      for (long k = rp[i] + lane; k < rp[i + 1]; k += L) s += v[k] * u[ci[k]];
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, L);
    if (ok && lane == 0) {
      q[i] = s;
      acc += u[i] * s;
    }
  }
  return acc;
}

// r = b - A x, p = M^-1 r; partials of b.b, r.r, r.p
template <int L>
__global__ __launch_bounds__(CSR_BS) void csr_resid_kernel(int n, int G, const long* rp,
                                                           const int* ci, const double* v,
                                                           const double* b, const double* x,
                                                           const double* dinv, double* r,
                                                           double* p, double* pa, double* pb,
                                                           double* pc) {
  __shared__ double lds[2 * (CSR_BS / 64) + 1];
  long r0, r1;
  csr_range(n, G, r0, r1);
  const int grp = threadIdx.x / L, lane = threadIdx.x % L;
  double bb = 0.0, rr = 0.0, rz = 0.0;
  for (long base = r0; base < r1; base += CSR_BS / L) {
    const long i = base + grp;
    const bool ok = i < r1;
    double s = 0.0;
    if (ok)
      for (long k = rp[i] + lane; k < rp[i + 1]; k += L) s += v[k] * x[ci[k]];
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, L);
    if (ok && lane == 0) {
      const double bi = b[i], ri = bi - s, zi = dinv[i] * ri;
      r[i] = ri;
      p[i] = zi;
      bb += bi * bi;
      rr += ri * ri;
      rz += ri * zi;
    }
  }
  const double sbb = block_sum<CSR_BS>(bb, lds);
  block_sum2<CSR_BS>(rr, rz, lds);
  if (threadIdx.x == 0) {
    pa[blockIdx.x] = sbb;
    pb[blockIdx.x] = rr;
    pc[blockIdx.x] = rz;
  }
}

// Eigen's start: b = 0 -> x = 0, done; |r|^2 below the threshold -> done
// with 0 iterations; else absNew = r.p
__global__ __launch_bounds__(CSR_BS) void csr_start_kernel(int n, int G, double tol2, double* x,
                                                           const double* pa, const double* pb,
                                                           const double* pc, CsrState* st) {
  __shared__ double lds[CSR_BS / 64 + 1];
  const double bb = csr_resum(pa, G, lds);
  const double rr = csr_resum(pb, G, lds);
  const double rz = csr_resum(pc, G, lds);
  const double thr = fmax(tol2 * bb, 2.2250738585072014e-308);
  if (bb == 0.0) {
    long r0, r1;
    csr_range(n, G, r0, r1);
    for (long i = r0 + threadIdx.x; i < r1; i += CSR_BS) x[i] = 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->bb = bb;
    st->thresh = thr;
    st->rr = bb == 0.0 ? 0.0 : rr;
    st->abs_new[0] = rz;
    st->iters = 0;
    st->done = (bb == 0.0 || rr < thr) ? 1 : 0;
  }
}

// q = A p, partials of p.q
template <int L>
__global__ __launch_bounds__(CSR_BS) void csr_spmv_kernel(int n, int G, const long* rp,
                                                          const int* ci, const double* v,
                                                          const double* p, double* q,
                                                          double* pa, const CsrState* st) {
  __shared__ double lds[CSR_BS / 64 + 1];
  __shared__ int flag;
  if (csr_done(st, &flag)) return;
  long r0, r1;
  csr_range(n, G, r0, r1);
  const double acc = csr_rows_spmv<L>(r0, r1, rp, ci, v, p, q);
  const double tot = block_sum<CSR_BS>(acc, lds);
  if (threadIdx.x == 0) pa[blockIdx.x] = tot;
}

// alpha = absNew / p.q; x += alpha p; r -= alpha q; z = M^-1 r; partials of
// r.r and r.z
__global__ __launch_bounds__(CSR_BS) void csr_update_kernel(int n, int G, int it, double* x,
                                                            double* r, double* z,
                                                            const double* p, const double* q,
                                                            const double* dinv,
                                                            const double* pa, double* pb,
                                                            double* pc, const CsrState* st) {
  __shared__ double lds[2 * (CSR_BS / 64) + 1];
  __shared__ int flag;
  if (csr_done(st, &flag)) return;
  const double pq = csr_resum(pa, G, lds);
  const double alpha = st->abs_new[it & 1] / pq;
  long r0, r1;
  csr_range(n, G, r0, r1);
  double rr = 0.0, rz = 0.0;
  for (long i = r0 + threadIdx.x; i < r1; i += CSR_BS) {
    const double pi = p[i], qi = q[i];
    x[i] += alpha * pi;
    const double ri = r[i] - alpha * qi;
    r[i] = ri;
    const double zi = dinv[i] * ri;
    z[i] = zi;
    rr += ri * ri;
    rz += ri * zi;
  }
  block_sum2<CSR_BS>(rr, rz, lds);
  if (threadIdx.x == 0) {
    pb[blockIdx.x] = rr;
    pc[blockIdx.x] = rz;
  }
}

// convergence test; else beta = r.z / absOld, p = z + beta p
__global__ __launch_bounds__(CSR_BS) void csr_dir_kernel(int n, int G, int it, const double* z,
                                                         double* p, const double* pb,
                                                         const double* pc, CsrState* st) {
  __shared__ double lds[CSR_BS / 64 + 1];
  __shared__ int flag;
  if (csr_done(st, &flag)) return;
  const double rr = csr_resum(pb, G, lds);
  if (rr < st->thresh) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->rr = rr;
      st->iters = it;
      st->done = 1;
    }
    return;
  }
  const double rz = csr_resum(pc, G, lds);
  const double beta = rz / st->abs_new[it & 1];
  long r0, r1;
  csr_range(n, G, r0, r1);
  for (long i = r0 + threadIdx.x; i < r1; i += CSR_BS) p[i] = z[i] + beta * p[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->abs_new[(it + 1) & 1] = rz;
    st->rr = rr;
    st->iters = it + 1;
  }
}

// partials of b.x
__global__ __launch_bounds__(CSR_BS) void csr_dot_kernel(int n, int G, const double* b,
                                                         const double* x, double* pa) {
  __shared__ double lds[CSR_BS / 64 + 1];
  long r0, r1;
  csr_range(n, G, r0, r1);
  double acc = 0.0;
  for (long i = r0 + threadIdx.x; i < r1; i += CSR_BS) acc += b[i] * x[i];
  const double tot = block_sum<CSR_BS>(acc, lds);
  if (threadIdx.x == 0) pa[blockIdx.x] = tot;
}

// lanes per row: the smallest power of two >= a quarter of the mean row
// length, 4..64; the grid: one workgroup per CSR_BS / L rows, at most CSR_GMAX
void csr_shape(long n, long nnz, int& L, int& G) {
  const double mean = n > 0 ? (double)nnz / (double)n : 0.0;
  L = 4;
  while (L < 64 && L < mean / 4.0) L *= 2;
  const long rows = CSR_BS / L;
  G = (int)std::min<long>(CSR_GMAX, std::max<long>(1, (n + rows - 1) / rows));
}

template <int L>
void launch_rows(hipStream_t s, int G, int n, bool resid, const long* rp, const int* ci,
                 const double* v, const double* b, const double* x, const CsrWork& w) {
  if (resid)
    hipLaunchKernelGGL(csr_resid_kernel<L>, dim3(G), dim3(CSR_BS), 0, s, n, G, rp, ci, v, b, x,
                       w.dinv, w.r, w.p, w.pa, w.pb, w.pc);
  else
    hipLaunchKernelGGL(csr_spmv_kernel<L>, dim3(G), dim3(CSR_BS), 0, s, n, G, rp, ci, v, w.p,
                       w.q, w.pa, w.st);
}

void launch_rows_any(hipStream_t s, int L, int G, int n, bool resid, const long* rp,
                     const int* ci, const double* v, const double* b, const double* x,
                     const CsrWork& w) {
  switch (L) {
    case 4: launch_rows<4>(s, G, n, resid, rp, ci, v, b, x, w); break;
    case 8: launch_rows<8>(s, G, n, resid, rp, ci, v, b, x, w); break;
    case 16: launch_rows<16>(s, G, n, resid, rp, ci, v, b, x, w); break;
    case 32: launch_rows<32>(s, G, n, resid, rp, ci, v, b, x, w); break;
    default: launch_rows<64>(s, G, n, resid, rp, ci, v, b, x, w); break;
  }
}

void read_state(hipStream_t s, const CsrWork& w, CsrState& h) {
  PGN_HIP(hipMemcpyAsync(&h, w.st, sizeof(CsrState), hipMemcpyDeviceToHost, s));
  PGN_HIP(hipStreamSynchronize(s));
}

// The solve over a device CSR matrix (full symmetric), device b and x (x is
// the warm start and the result); `work` holds csr_work_bytes(n) bytes.
// normA: |A|_F over the caller's stored entries when the host computed it
// (a lower layout), < 0 to compute it here over the device matrix.  shifts:
// the cumulative diagonal shifts applied, in order (for the host entry
// point's write-back).  Allocates nothing.
void csr_solve_device(hipStream_t s, int n, long nnz, const long* rp, const int* ci, double* v,
                      const double* b, double* x, const pgn_linear_opts& o, double normA,
                      char* work, pgn_linear_result& res, std::vector<double>* shifts) {
  res = pgn_linear_result{};
  res.status = 1;
  if (n == 0) {
    res.status = 0;
    return;
  }
  int L, G;
  csr_shape(n, nnz, L, G);
  const CsrWork w = csr_carve(work, n);
  hipLaunchKernelGGL(csr_state_init_kernel, dim3(1), dim3(64), 0, s, w.st);
  hipLaunchKernelGGL(csr_setup_kernel, dim3(G), dim3(CSR_BS), 0, s, n, nnz, G, rp, ci, v,
                     w.dpos, w.pa, w.st);
  hipLaunchKernelGGL(csr_sum_kernel, dim3(1), dim3(CSR_BS), 0, s, w.pa, G, &w.st->normA2);
  PGN_HIP(hipGetLastError());
  CsrState h;
  read_state(s, w, h);
  PGN_CHECK(!h.bad, "malformed CSR matrix: row pointers must run from 0 to nnz without "
                    "decreasing and column indices must lie in [0, n)");
  PGN_CHECK(!h.nodiag, "every row must store its diagonal entry");
  // LinearSolver.cpp:57-66: normA = mA.norm() over the stored entries
  const double regT = 1e-6 * (normA >= 0.0 ? normA : std::sqrt(h.normA2));
  hipLaunchKernelGGL(csr_copy_kernel, dim3(G), dim3(CSR_BS), 0, s, n, G, x, w.xprev);
  const double tol2 = o.tol * o.tol;
  for (int i = 0; i < o.reg_iters; ++i) {
    if (i != 0) {
      // LinearSolver_EigenCG.cpp:73: mA += R regT 10^(i-1), cumulative
      const double sh = regT * std::pow(10.0, i - 1);
      hipLaunchKernelGGL(csr_shift_kernel, dim3(G), dim3(CSR_BS), 0, s, n, G, w.dpos, v, sh);
      if (shifts) shifts->push_back(sh);
      res.reg_used = i;
    }
    hipLaunchKernelGGL(csr_restore_kernel, dim3(G), dim3(CSR_BS), 0, s, n, G, w.xprev, x,
                       w.dpos, v, w.dinv, w.st);
    launch_rows_any(s, L, G, n, true, rp, ci, v, b, x, w);
    hipLaunchKernelGGL(csr_start_kernel, dim3(G), dim3(CSR_BS), 0, s, n, G, tol2, x, w.pa,
                       w.pb, w.pc, w.st);
    PGN_HIP(hipGetLastError());
    int it = 0, chunk = 8;
    while (it < o.max_iters) {
      const int k = std::min(chunk, o.max_iters - it);
      for (int j = 0; j < k; ++j, ++it) {
        launch_rows_any(s, L, G, n, false, rp, ci, v, b, x, w);
        hipLaunchKernelGGL(csr_update_kernel, dim3(G), dim3(CSR_BS), 0, s, n, G, it, x, w.r,
                           w.z, w.p, w.q, w.dinv, w.pa, w.pb, w.pc, w.st);
        hipLaunchKernelGGL(csr_dir_kernel, dim3(G), dim3(CSR_BS), 0, s, n, G, it, w.z, w.p,
                           w.pb, w.pc, w.st);
      }
      PGN_HIP(hipGetLastError());
      read_state(s, w, h);
      if (h.done) break;
      chunk = std::min(chunk * 2, 64);
    }
    hipLaunchKernelGGL(csr_dot_kernel, dim3(G), dim3(CSR_BS), 0, s, n, G, b, x, w.pa);
    hipLaunchKernelGGL(csr_sum_kernel, dim3(1), dim3(CSR_BS), 0, s, w.pa, G, &w.st->bx);
    PGN_HIP(hipGetLastError());
    read_state(s, w, h);
    res.iterations += h.iters;
    res.rel_residual = h.bb > 0.0 ? std::sqrt(h.rr / h.bb) : 0.0;
    if (h.bx < 0.0) continue;  // LinearSolver_EigenCG.cpp:97-101: indefinite
    res.status = 0;
    break;
  }
}

template <class F>
int cguard(F&& f) {
  try {
    f();
    return PGN_OK;
  } catch (const Error& e) {
    return api_guard_set(e.what(), e.code);
  } catch (const std::bad_alloc&) {
    return api_guard_set("host allocation failed", PGN_ERR_NOMEM);
  } catch (const std::exception& e) {
    return api_guard_set(e.what(), PGN_ERR_INVALID);
  }
}

void check_opts(const pgn_linear_opts& o) {
  PGN_CHECK(o.solver == PGN_LS_EIGENCG, "the CSR solve is LinearSolver_EigenCG (PGN_LS_EIGENCG)");
  PGN_CHECK(o.precond == 0, "the CSR solve preconditions with Jacobi (precond 0)");
  PGN_CHECK(o.tol >= 0.0 && o.max_iters >= 0 && o.reg_iters >= 1,
            "tol >= 0, max_iters >= 0 and reg_iters >= 1");
}

}  // namespace

namespace pgn {
// Once per device, with the device current: the address of that device's
// copy of `scratch` (called from pgn_ctx_create).
void csr_scratch_setup(int device) {
  PGN_CHECK(device >= 0 && device < CSR_MAX_DEVICES, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* ptr = nullptr;
  PGN_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  PGN_CHECK(ptr != nullptr, "no device address for the CSR scratch array");
  g_scratch_ptr[device] = ptr;
}
}  // namespace pgn

extern "C" {

int pgn_csr_linear_solve(pgn_ctx* ctx, int n, int layout, const long* rowptr, const int* col,
                         double* val, const double* b, double* x, const pgn_linear_opts* opts,
                         pgn_linear_result* res) {
  return cguard([&] {
    PGN_CHECK(ctx && rowptr && b && x && res, "null argument");
    PGN_CHECK(n >= 0, "n must be >= 0");
    PGN_CHECK(layout == PGN_LAYOUT_LOWER || layout == PGN_LAYOUT_FULL,
              "layout must be PGN_LAYOUT_LOWER or PGN_LAYOUT_FULL");
    pgn_linear_opts o;
    pgn_linear_default_opts(&o);
    if (opts) o = *opts;
    check_opts(o);
    const long nnz_in = rowptr[n];
    PGN_CHECK(rowptr[0] == 0 && nnz_in >= 0, "rowptr must start at 0");
    PGN_CHECK(nnz_in == 0 || (col && val), "null argument");
    // host checks: the device never sees an out-of-range index
    std::vector<long> dpos(n, -1);
    for (int i = 0; i < n; ++i) {
      PGN_CHECK(rowptr[i + 1] >= rowptr[i] && rowptr[i + 1] <= nnz_in,
                "row pointers must not decrease and must end at nnz");
      for (long k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        PGN_CHECK(col[k] >= 0 && col[k] < n, "column index out of range");
        if (col[k] == i && dpos[i] < 0) dpos[i] = k;
      }
      PGN_CHECK(dpos[i] >= 0, "every row must store its diagonal entry");
    }
    // GetMatrixData over the caller's stored entries, in their order
    double normA = -1.0;
    if (layout == PGN_LAYOUT_LOWER) {
      double a2 = 0.0;
      for (long k = 0; k < nnz_in; ++k) a2 += val[k] * val[k];
      normA = std::sqrt(a2);
    }
    // the full symmetric matrix: the lower triangle mirrored (entries above
    // the diagonal of a lower layout, such as the upper halves of
    // AMatrixSd's whole diagonal blocks, are read by nothing, as with
    // selfadjointView<Lower>), columns sorted
    std::vector<long> rp;
    std::vector<int> ci;
    std::vector<double> v;
    if (layout == PGN_LAYOUT_FULL) {
      rp.assign(rowptr, rowptr + n + 1);
      ci.assign(col, col + nnz_in);
      v.assign(val, val + nnz_in);
    } else {
      std::vector<long> cnt(n + 1, 0);
      for (int i = 0; i < n; ++i)
        for (long k = rowptr[i]; k < rowptr[i + 1]; ++k) {
          if (col[k] > i) continue;
          cnt[i + 1]++;
          if (col[k] != i) cnt[col[k] + 1]++;
        }
      for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
      rp = cnt;
      ci.resize(rp[n]);
      v.resize(rp[n]);
      std::vector<long> fill(rp.begin(), rp.end() - 1);
      for (int i = 0; i < n; ++i)  // row i ascending: every (j, i), j < i, lands in order
        for (long k = rowptr[i]; k < rowptr[i + 1]; ++k) {
          const int c = col[k];
          if (c > i) continue;
          ci[fill[i]] = c;
          v[fill[i]++] = val[k];
          if (c != i) {
            ci[fill[c]] = i;
            v[fill[c]++] = val[k];
          }
        }
      for (int i = 0; i < n; ++i) {  // sort each row by column
        std::vector<std::pair<int, double>> row;
        for (long k = rp[i]; k < rp[i + 1]; ++k) row.emplace_back(ci[k], v[k]);
        std::stable_sort(row.begin(), row.end(),
                         [](const auto& a, const auto& c) { return a.first < c.first; });
        for (long k = rp[i]; k < rp[i + 1]; ++k) {
          ci[k] = row[k - rp[i]].first;
          v[k] = row[k - rp[i]].second;
        }
      }
    }
    PGN_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    DBuf<long> drp;
    DBuf<int> dci;
    DBuf<double> dv, db, dx;
    DBuf<char> work;
    drp.upload(rp, s);
    dci.upload(ci, s);
    dv.upload(v, s);
    db.upload(b, n, s);
    dx.upload(x, n, s);
    work.alloc(csr_work_bytes(std::max(n, 1)));
    std::vector<double> shifts;
    pgn_linear_result r;
    csr_solve_device(s, n, rp[n], drp.p, dci.p, dv.p, db.p, dx.p, o, normA, work.p, r, &shifts);
    dx.download(x, n, s);
    PGN_HIP(hipStreamSynchronize(s));
    // the shifts stay in the caller's matrix, added in the same order
    for (double sh : shifts)
      for (int i = 0; i < n; ++i) val[dpos[i]] += sh;
    *res = r;
  });
}

int pgn_csr_linear_solve_noalloc(pgn_ctx* ctx, int n, const long* rowptr, const int* col,
                                 double* val, const double* b, double* x,
                                 const pgn_linear_opts* opts, pgn_linear_result* res) {
  return cguard([&] {
    PGN_CHECK(ctx && rowptr && col && val && b && x && res, "null argument");
    PGN_CHECK(n >= 0, "n must be >= 0");
    PGN_CHECK(n <= PGN_CSR_NOALLOC_MAX_ROWS,
              "pgn_csr_linear_solve_noalloc: n exceeds PGN_CSR_NOALLOC_MAX_ROWS (262144 rows)");
    pgn_linear_opts o;
    pgn_linear_default_opts(&o);
    if (opts) o = *opts;
    check_opts(o);
    PGN_HIP(hipSetDevice(ctx->device));
    void* base = nullptr;
    {
      std::lock_guard<std::mutex> lk(g_scratch_mu);
      base = ctx->device < CSR_MAX_DEVICES ? g_scratch_ptr[ctx->device] : nullptr;
    }
    PGN_CHECK(base, "no scratch address for this device (context not created by pgn_ctx_create)");
    hipStream_t s = ctx->stream;
    long nnz = 0;
    PGN_HIP(hipMemcpyAsync(&nnz, rowptr + n, sizeof(long), hipMemcpyDeviceToHost, s));
    PGN_HIP(hipStreamSynchronize(s));
    PGN_CHECK(nnz >= 0, "malformed CSR matrix: negative nnz");
    pgn_linear_result r;
    csr_solve_device(s, n, nnz, rowptr, col, val, b, x, o, -1.0, (char*)base, r, nullptr);
    *res = r;
  });
}

}  // extern "C"
