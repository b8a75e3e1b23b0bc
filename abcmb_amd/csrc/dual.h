// Forward-mode dual numbers for the tangent-linear kernels.
//
// D<N> carries a value and its derivatives along N parameter directions.
// D<0> is a plain double in a struct, so one templated code path serves the
// primal kernels (N = 0) and the tangent kernels (N > 0).  Only the
// operations the perturbation / line-of-sight / lensing code uses are
// provided; comparisons and branches look at the value.
#pragma once
#include <hip/hip_runtime.h>

namespace abcmb {

template <int N>
struct D {
  double v;
  double d[N];
  __device__ __forceinline__ D() {}
  __device__ __forceinline__ D(double x) : v(x) {
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = 0.0;
  }
};

template <>
struct D<0> {
  double v;
  __device__ __forceinline__ D() {}
  __device__ __forceinline__ D(double x) : v(x) {}
};

#define ABCMB_DLOOP(N) _Pragma("unroll") for (int i_ = 0; i_ < (N); ++i_)

template <int N>
__device__ __forceinline__ double dget(const D<N>& a, int i) {
  if constexpr (N > 0) return a.d[i];
  else return 0.0;
}

template <int N>
__device__ __forceinline__ D<N> operator+(const D<N>& a, const D<N>& b) {
  D<N> r; r.v = a.v + b.v;
  if constexpr (N > 0) { ABCMB_DLOOP(N) r.d[i_] = a.d[i_] + b.d[i_]; }
  return r;
}
template <int N>
__device__ __forceinline__ D<N> operator-(const D<N>& a, const D<N>& b) {
  D<N> r; r.v = a.v - b.v;
  if constexpr (N > 0) { ABCMB_DLOOP(N) r.d[i_] = a.d[i_] - b.d[i_]; }
  return r;
}
template <int N>
__device__ __forceinline__ D<N> operator-(const D<N>& a) {
  D<N> r; r.v = -a.v;
  if constexpr (N > 0) { ABCMB_DLOOP(N) r.d[i_] = -a.d[i_]; }
  return r;
}
template <int N>
__device__ __forceinline__ D<N> operator*(const D<N>& a, const D<N>& b) {
  D<N> r; r.v = a.v * b.v;
  if constexpr (N > 0) { ABCMB_DLOOP(N) r.d[i_] = a.d[i_] * b.v + a.v * b.d[i_]; }
  return r;
}
template <int N>
__device__ __forceinline__ D<N> operator/(const D<N>& a, const D<N>& b) {
  D<N> r; r.v = a.v / b.v;
  if constexpr (N > 0) {
    double ib = 1.0 / b.v;
    ABCMB_DLOOP(N) r.d[i_] = (a.d[i_] - r.v * b.d[i_]) * ib;
  }
  return r;
}
// mixed with plain doubles
template <int N> __device__ __forceinline__ D<N> operator+(const D<N>& a, double b) { D<N> r = a; r.v += b; return r; }
template <int N> __device__ __forceinline__ D<N> operator+(double b, const D<N>& a) { D<N> r = a; r.v += b; return r; }
template <int N> __device__ __forceinline__ D<N> operator-(const D<N>& a, double b) { D<N> r = a; r.v -= b; return r; }
template <int N> __device__ __forceinline__ D<N> operator-(double b, const D<N>& a) {
  D<N> r; r.v = b - a.v;
  if constexpr (N > 0) { ABCMB_DLOOP(N) r.d[i_] = -a.d[i_]; }
  return r;
}
template <int N> __device__ __forceinline__ D<N> operator*(const D<N>& a, double b) {
  D<N> r; r.v = a.v * b;
  if constexpr (N > 0) { ABCMB_DLOOP(N) r.d[i_] = a.d[i_] * b; }
  return r;
}
template <int N> __device__ __forceinline__ D<N> operator*(double b, const D<N>& a) { return a * b; }
template <int N> __device__ __forceinline__ D<N> operator/(const D<N>& a, double b) {
  D<N> r; r.v = a.v / b;
  if constexpr (N > 0) {
    double ib = 1.0 / b;
    ABCMB_DLOOP(N) r.d[i_] = a.d[i_] * ib;
  }
  return r;
}
template <int N> __device__ __forceinline__ D<N> operator/(double b, const D<N>& a) {
  D<N> r; r.v = b / a.v;
  if constexpr (N > 0) {
    double s = -r.v / a.v;
    ABCMB_DLOOP(N) r.d[i_] = s * a.d[i_];
  }
  return r;
}
// 1 / a with one division: the value's reciprocal, and -d / a^2 from it
template <int N> __device__ __forceinline__ D<N> dinv(const D<N>& a) {
  D<N> r; r.v = 1.0 / a.v;
  if constexpr (N > 0) {
    const double s = -r.v * r.v;
    ABCMB_DLOOP(N) r.d[i_] = s * a.d[i_];
  }
  return r;
}
template <int N> __device__ __forceinline__ D<N>& operator+=(D<N>& a, const D<N>& b) { a = a + b; return a; }
template <int N> __device__ __forceinline__ D<N>& operator-=(D<N>& a, const D<N>& b) { a = a - b; return a; }

// a * b + c with a plain coefficient a (the common mat-vec step)
template <int N>
__device__ __forceinline__ D<N> fma_c(double a, const D<N>& b, const D<N>& c) {
  D<N> r; r.v = fma(a, b.v, c.v);
  if constexpr (N > 0) { ABCMB_DLOOP(N) r.d[i_] = fma(a, b.d[i_], c.d[i_]); }
  return r;
}

// elementary functions: f(a) with f'(a) chained into the derivatives
template <int N>
__device__ __forceinline__ D<N> chain(const D<N>& a, double fv, double fp) {
  D<N> r; r.v = fv;
  if constexpr (N > 0) { ABCMB_DLOOP(N) r.d[i_] = fp * a.d[i_]; }
  return r;
}
template <int N> __device__ __forceinline__ D<N> dexp(const D<N>& a) { double e = exp(a.v); return chain(a, e, e); }
template <int N> __device__ __forceinline__ D<N> dlog(const D<N>& a) { return chain(a, log(a.v), 1.0 / a.v); }
template <int N> __device__ __forceinline__ D<N> dsqrt(const D<N>& a) {
  double s = sqrt(a.v);
  return chain(a, s, 0.5 / s);
}
template <int N> __device__ __forceinline__ D<N> dpow(const D<N>& a, double p) {
  double s = pow(a.v, p);
  return chain(a, s, p * pow(a.v, p - 1.0));
}
template <int N> __device__ __forceinline__ D<N> dcos(const D<N>& a) { return chain(a, cos(a.v), -sin(a.v)); }
template <int N> __device__ __forceinline__ D<N> dtanh(const D<N>& a) {
  const double t = tanh(a.v);
  return chain(a, t, 1.0 - t * t);
}
template <int N> __device__ __forceinline__ D<N> dasin(const D<N>& a) {
  return chain(a, asin(a.v), 1.0 / sqrt(1.0 - a.v * a.v));
}

}  // namespace abcmb
