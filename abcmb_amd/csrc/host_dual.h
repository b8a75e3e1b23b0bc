// Forward-mode dual numbers for the host producer.
//
// Dual<N>: a value and its derivatives along N parameter directions, carried
// through the background integrals, the recombination ODEs and the
// reionization root so that the tangent seeds of the perturbation kernels
// are exact derivatives of the producer (abcmb_host_background_tangent).
// Dl<S>: a value and one derivative with respect to a local variable (the
// state or the time), over either scalar type; HyRex's post-Saha expansions
// take such derivatives (jax.grad in abcmb/hyrex/helium.py:297-373,
// hydrogen.py:190-276).
//
// Every operation computes the value exactly as the plain double code does,
// so the value part of a dual evaluation equals the double evaluation.
#pragma once
#include <cmath>
#include <type_traits>

namespace abcmb_ad {

template <int N>
struct Dual {
  double v;
  double d[N];
  Dual() : v(0.0) {
    for (int i = 0; i < N; ++i) d[i] = 0.0;
  }
  Dual(double x) : v(x) {  // NOLINT: a constant
    for (int i = 0; i < N; ++i) d[i] = 0.0;
  }
};

template <int N> inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v + b.v);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N> inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v - b.v);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N> inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v * b.v);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N> inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v / b.v);
  const double ib = 1.0 / b.v;
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <int N> inline Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r(-a.v);
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N> inline Dual<N> operator+(const Dual<N>& a, double b) {
  Dual<N> r = a;
  r.v = a.v + b;
  return r;
}
template <int N> inline Dual<N> operator+(double b, const Dual<N>& a) {
  Dual<N> r = a;
  r.v = b + a.v;
  return r;
}
template <int N> inline Dual<N> operator-(const Dual<N>& a, double b) {
  Dual<N> r = a;
  r.v = a.v - b;
  return r;
}
template <int N> inline Dual<N> operator-(double b, const Dual<N>& a) {
  Dual<N> r(b - a.v);
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N> inline Dual<N> operator*(const Dual<N>& a, double b) {
  Dual<N> r(a.v * b);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
  return r;
}
template <int N> inline Dual<N> operator*(double b, const Dual<N>& a) {
  Dual<N> r(b * a.v);
  for (int i = 0; i < N; ++i) r.d[i] = b * a.d[i];
  return r;
}
template <int N> inline Dual<N> operator/(const Dual<N>& a, double b) {
  Dual<N> r(a.v / b);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / b;
  return r;
}
template <int N> inline Dual<N> operator/(double b, const Dual<N>& a) {
  Dual<N> r(b / a.v);
  const double s = -r.v / a.v;
  for (int i = 0; i < N; ++i) r.d[i] = s * a.d[i];
  return r;
}
template <int N, class U> inline Dual<N>& operator+=(Dual<N>& a, const U& b) { return a = a + b; }
template <int N, class U> inline Dual<N>& operator-=(Dual<N>& a, const U& b) { return a = a - b; }
template <int N, class U> inline Dual<N>& operator*=(Dual<N>& a, const U& b) { return a = a * b; }

template <int N> inline Dual<N> chain(const Dual<N>& a, double f, double df) {
  Dual<N> r(f);
  for (int i = 0; i < N; ++i) r.d[i] = df * a.d[i];
  return r;
}
template <int N> inline Dual<N> exp(const Dual<N>& a) {
  const double e = std::exp(a.v);
  return chain(a, e, e);
}
template <int N> inline Dual<N> log(const Dual<N>& a) { return chain(a, std::log(a.v), 1.0 / a.v); }
template <int N> inline Dual<N> sqrt(const Dual<N>& a) {
  const double s = std::sqrt(a.v);
  return chain(a, s, 0.5 / s);
}
template <int N> inline Dual<N> pow(const Dual<N>& a, double p) {
  return chain(a, std::pow(a.v, p), p * std::pow(a.v, p - 1.0));
}
template <int N> inline Dual<N> cos(const Dual<N>& a) { return chain(a, std::cos(a.v), -std::sin(a.v)); }
template <int N> inline Dual<N> tanh(const Dual<N>& a) {
  const double t = std::tanh(a.v);
  return chain(a, t, 1.0 - t * t);
}

// value and one local derivative, over double or Dual<N>
template <class S>
struct Dl {
  S v, d;
  Dl() : v(0.0), d(0.0) {}
  Dl(const S& x, const S& dx = S(0.0)) : v(x), d(dx) {}  // NOLINT
};

template <class T> struct is_dl : std::false_type {};
template <class S> struct is_dl<Dl<S>> : std::true_type {};
template <class U> using not_dl = std::enable_if_t<!is_dl<U>::value, int>;

template <class S> inline Dl<S> operator+(const Dl<S>& a, const Dl<S>& b) { return Dl<S>(a.v + b.v, a.d + b.d); }
template <class S> inline Dl<S> operator-(const Dl<S>& a, const Dl<S>& b) { return Dl<S>(a.v - b.v, a.d - b.d); }
template <class S> inline Dl<S> operator*(const Dl<S>& a, const Dl<S>& b) {
  return Dl<S>(a.v * b.v, a.d * b.v + a.v * b.d);
}
template <class S> inline Dl<S> operator/(const Dl<S>& a, const Dl<S>& b) {
  return Dl<S>(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v));
}
template <class S> inline Dl<S> operator-(const Dl<S>& a) { return Dl<S>(-a.v, -a.d); }
template <class S, class U, not_dl<U> = 0> inline Dl<S> operator+(const Dl<S>& a, const U& b) {
  return Dl<S>(a.v + b, a.d);
}
template <class S, class U, not_dl<U> = 0> inline Dl<S> operator+(const U& b, const Dl<S>& a) {
  return Dl<S>(a.v + b, a.d);
}
template <class S, class U, not_dl<U> = 0> inline Dl<S> operator-(const Dl<S>& a, const U& b) {
  return Dl<S>(a.v - b, a.d);
}
template <class S, class U, not_dl<U> = 0> inline Dl<S> operator-(const U& b, const Dl<S>& a) {
  return Dl<S>(b - a.v, -a.d);
}
template <class S, class U, not_dl<U> = 0> inline Dl<S> operator*(const Dl<S>& a, const U& b) {
  return Dl<S>(a.v * b, a.d * b);
}
template <class S, class U, not_dl<U> = 0> inline Dl<S> operator*(const U& b, const Dl<S>& a) {
  return Dl<S>(a.v * b, a.d * b);
}
template <class S, class U, not_dl<U> = 0> inline Dl<S> operator/(const Dl<S>& a, const U& b) {
  return Dl<S>(a.v / b, a.d / b);
}
template <class S, class U, not_dl<U> = 0> inline Dl<S> operator/(const U& b, const Dl<S>& a) {
  return Dl<S>(b / a.v, -b * a.d / (a.v * a.v));
}
template <class S> inline Dl<S> exp(const Dl<S>& a) {
  using std::exp;
  const S e = exp(a.v);
  return Dl<S>(e, e * a.d);
}
template <class S> inline Dl<S> log(const Dl<S>& a) {
  using std::log;
  return Dl<S>(log(a.v), a.d / a.v);
}
template <class S> inline Dl<S> sqrt(const Dl<S>& a) {
  using std::sqrt;
  const S s = sqrt(a.v);
  return Dl<S>(s, 0.5 * a.d / s);
}
template <class S> inline Dl<S> tanh(const Dl<S>& a) {
  using std::tanh;
  const S t = tanh(a.v);
  return Dl<S>(t, (1.0 - t * t) * a.d);
}
template <class S> inline Dl<S> pow(const Dl<S>& a, double p) {
  using std::pow;
  const S s = pow(a.v, p);
  return Dl<S>(s, p * pow(a.v, p - 1.0) * a.d);
}

// the plain double value of any of the scalar types
inline double dbl(double x) { return x; }
template <int N> inline double dbl(const Dual<N>& x) { return x.v; }
template <class S> inline double dbl(const Dl<S>& x) { return dbl(x.v); }

}  // namespace abcmb_ad
