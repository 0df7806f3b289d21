// reduce.hpp -- deterministic workgroup reductions for wave64 (gfx950).
#pragma once

#include <hip/hip_runtime.h>

namespace pgn {

// DPP row shift right by N within each 16-lane row (lanes shifted in from
// outside the row read 0): v[lane - N] or 0
template <int N, class T>
__device__ __forceinline__ T dpp_shr(T v) {
  return __builtin_amdgcn_update_dpp(T(0), v, 0x110 + N, 0xF, 0xF, true);
}

// v of one lane as a wavefront-uniform value
__device__ __forceinline__ double read_lane(double v, int lane) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));
}

// Sum over the wavefront (all 64 lanes active), in every lane: inclusive
// scans within the four 16-lane rows by DPP shifts leave the row totals in
// lanes 15, 31, 47, 63, added in that fixed order.  DPP moves take a few
// cycles where shuffles go through LDS permutes.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_shr<1>(v);
  v += dpp_shr<2>(v);
  v += dpp_shr<4>(v);
  v += dpp_shr<8>(v);
  return (read_lane(v, 15) + read_lane(v, 31)) + (read_lane(v, 47) + read_lane(v, 63));
}

// Sum over the workgroup in a fixed order; the result is valid in thread 0.
template <int BS>
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < BS / 64; ++i) r += lds[i];
  }
  __syncthreads();
  return r;
}

// Two sums at once (fixed order), valid in thread 0.
template <int BS>
__device__ __forceinline__ void block_sum2(double& a, double& b, double* lds) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    lds[w] = a;
    lds[BS / 64 + w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ra = 0.0, rb = 0.0;
#pragma unroll
    for (int i = 0; i < BS / 64; ++i) {
      ra += lds[i];
      rb += lds[BS / 64 + i];
    }
    a = ra;
    b = rb;
  }
  __syncthreads();
}

// Thread 0 of every workgroup calls this after writing its partial result.
// Returns true in the workgroup that arrives last (for all its threads); that
// workgroup then sees every other workgroup's partials.  The ticket is reset
// for the next launch.
__device__ __forceinline__ bool last_block(unsigned* ticket, bool* flag_lds) {
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned t = atomicAdd(ticket, 1u);
    *flag_lds = (t == gridDim.x - 1);
  }
  __syncthreads();
  const bool last = *flag_lds;
  if (last) {
    __threadfence();
    if (threadIdx.x == 0) *ticket = 0u;
  }
  return last;
}

// Sum `count` doubles (stride `stride`) in a fixed order by one workgroup;
// valid in thread 0.
template <int BS>
__device__ __forceinline__ double block_sum_array(const double* a, int count,
                                                  int stride, double* lds) {
  double v = 0.0;
  for (int i = threadIdx.x; i < count; i += BS)
    v += __hip_atomic_load(const_cast<double*>(a) + (long)i * stride, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  return block_sum<BS>(v, lds);
}

}  // namespace pgn
