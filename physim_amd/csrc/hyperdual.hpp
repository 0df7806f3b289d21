// hyperdual.hpp -- forward-mode automatic differentiation numbers for the
// element kernels: value, gradient over N variables and (when H is true) the
// lower triangle of the Hessian, carried through every arithmetic operation.
// Used where an element energy is too involved for hand-derived second
// derivatives (the rod bend-twist element: parallel transport, Rodrigues
// rotations and an atan2).
#pragma once

#include <hip/hip_runtime.h>

namespace pgn {

template <class T, int N, bool H>
struct HD {
  static constexpr int NH = H ? N * (N + 1) / 2 : 1;
  T v;
  T g[N];
  T h[NH];
};

template <class T, int N, bool H>
__host__ __device__ __forceinline__ HD<T, N, H> hd_const(T c) {
  HD<T, N, H> r;
  r.v = c;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = T(0);
#pragma unroll
  for (int i = 0; i < HD<T, N, H>::NH; ++i) r.h[i] = T(0);
  return r;
}

template <class T, int N, bool H>
__host__ __device__ __forceinline__ HD<T, N, H> hd_var(T c, int k) {
  HD<T, N, H> r = hd_const<T, N, H>(c);
  r.g[k] = T(1);
  return r;
}

template <class T, int N, bool H>
__host__ __device__ __forceinline__ HD<T, N, H> operator+(const HD<T, N, H>& a,
                                                           const HD<T, N, H>& b) {
  HD<T, N, H> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] + b.g[i];
  if (H) {
#pragma unroll
    for (int i = 0; i < HD<T, N, H>::NH; ++i) r.h[i] = a.h[i] + b.h[i];
  }
  return r;
}

template <class T, int N, bool H>
__host__ __device__ __forceinline__ HD<T, N, H> operator-(const HD<T, N, H>& a,
                                                           const HD<T, N, H>& b) {
  HD<T, N, H> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] - b.g[i];
  if (H) {
#pragma unroll
    for (int i = 0; i < HD<T, N, H>::NH; ++i) r.h[i] = a.h[i] - b.h[i];
  }
  return r;
}

template <class T, int N, bool H>
__host__ __device__ __forceinline__ HD<T, N, H> operator*(const HD<T, N, H>& a, T s) {
  HD<T, N, H> r;
  r.v = a.v * s;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] * s;
  if (H) {
#pragma unroll
    for (int i = 0; i < HD<T, N, H>::NH; ++i) r.h[i] = a.h[i] * s;
  }
  return r;
}

template <class T, int N, bool H>
__host__ __device__ __forceinline__ HD<T, N, H> operator+(const HD<T, N, H>& a, T c) {
  HD<T, N, H> r = a;
  r.v += c;
  return r;
}

template <class T, int N, bool H>
__host__ __device__ __forceinline__ HD<T, N, H> operator*(const HD<T, N, H>& a,
                                                           const HD<T, N, H>& b) {
  HD<T, N, H> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = a.v * b.g[i] + b.v * a.g[i];
  if (H) {
    int k = 0;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j, ++k)
        r.h[k] = a.v * b.h[k] + b.v * a.h[k] + a.g[i] * b.g[j] + a.g[j] * b.g[i];
  }
  return r;
}

// f(a) given f, f' and f'' at a.v
template <class T, int N, bool H>
__host__ __device__ __forceinline__ HD<T, N, H> hd_chain(const HD<T, N, H>& a, T f, T f1,
                                                          T f2) {
  HD<T, N, H> r;
  r.v = f;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = f1 * a.g[i];
  if (H) {
    int k = 0;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j, ++k) r.h[k] = f1 * a.h[k] + f2 * a.g[i] * a.g[j];
  }
  return r;
}

template <class T, int N, bool H>
__host__ __device__ __forceinline__ HD<T, N, H> hd_recip(const HD<T, N, H>& a) {
  const T iv = T(1) / a.v;
  return hd_chain(a, iv, -iv * iv, T(2) * iv * iv * iv);
}

template <class T, int N, bool H>
__host__ __device__ __forceinline__ HD<T, N, H> hd_sqrt(const HD<T, N, H>& a) {
  const T s = sqrt(a.v);
  return hd_chain(a, s, T(0.5) / s, T(-0.25) / (a.v * s));
}

template <class T, int N, bool H>
__host__ __device__ __forceinline__ HD<T, N, H> hd_sin(const HD<T, N, H>& a) {
  const T s = sin(a.v), c = cos(a.v);
  return hd_chain(a, s, c, -s);
}

template <class T, int N, bool H>
__host__ __device__ __forceinline__ HD<T, N, H> hd_cos(const HD<T, N, H>& a) {
  const T s = sin(a.v), c = cos(a.v);
  return hd_chain(a, c, -s, -c);
}

// atan2(y, x): first derivative (x dy - y dx) / r^2, second derivative of the
// same quotient, symmetrized.
template <class T, int N, bool H>
__host__ __device__ __forceinline__ HD<T, N, H> hd_atan2(const HD<T, N, H>& y,
                                                          const HD<T, N, H>& x) {
  HD<T, N, H> r;
  const T r2 = x.v * x.v + y.v * y.v;
  const T ir2 = T(1) / r2;
  r.v = atan2(y.v, x.v);
  T w[N], dr[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    w[i] = x.v * y.g[i] - y.v * x.g[i];
    dr[i] = T(2) * (x.v * x.g[i] + y.v * y.g[i]);
    r.g[i] = w[i] * ir2;
  }
  if (H) {
    int k = 0;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j, ++k) {
        const T sym = x.v * y.h[k] - y.v * x.h[k];  // the antisymmetric parts cancel
        r.h[k] = sym * ir2 - T(0.5) * (w[i] * dr[j] + w[j] * dr[i]) * ir2 * ir2;
      }
  }
  return r;
}

template <class S>
struct V3 {
  S c[3];
};

template <class S>
__host__ __device__ __forceinline__ S dot3(const V3<S>& a, const V3<S>& b) {
  return a.c[0] * b.c[0] + a.c[1] * b.c[1] + a.c[2] * b.c[2];
}

template <class S>
__host__ __device__ __forceinline__ V3<S> cross3(const V3<S>& a, const V3<S>& b) {
  V3<S> r;
  r.c[0] = a.c[1] * b.c[2] - a.c[2] * b.c[1];
  r.c[1] = a.c[2] * b.c[0] - a.c[0] * b.c[2];
  r.c[2] = a.c[0] * b.c[1] - a.c[1] * b.c[0];
  return r;
}

template <class S>
__host__ __device__ __forceinline__ V3<S> add3(const V3<S>& a, const V3<S>& b) {
  V3<S> r;
#pragma unroll
  for (int i = 0; i < 3; ++i) r.c[i] = a.c[i] + b.c[i];
  return r;
}

template <class S, class F>
__host__ __device__ __forceinline__ V3<S> scale3(const V3<S>& a, const F& s) {
  V3<S> r;
#pragma unroll
  for (int i = 0; i < 3; ++i) r.c[i] = a.c[i] * s;
  return r;
}

}  // namespace pgn
