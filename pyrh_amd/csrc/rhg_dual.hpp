// Forward-mode dual numbers for the response functions: a value and one
// directional derivative, with the arithmetic and the elementary functions
// the opacity and formal-solution code uses.  Comparisons and branches act
// on the value, so a dual evaluation follows the same branch as the plain
// one and returns the exact derivative of what the plain code computes.
#pragma once

#include <type_traits>

#include "rhg_device.hpp"

namespace rhg {

// the plain overloads stay visible beside the dual ones below
using ::fabs;
using ::sqrt;

template <typename R>
struct Dual {
  R v, d;
  Dual() = default;
  template <typename S, typename = std::enable_if_t<std::is_arithmetic<S>::value>>
  __host__ __device__ __forceinline__ constexpr Dual(S val) : v(R(val)), d(R(0)) {}
  __host__ __device__ __forceinline__ constexpr Dual(R val, R der) : v(val), d(der) {}
  // between precisions (the FP32 form's Voigt functions inside FP64 points)
  template <typename S, typename = std::enable_if_t<!std::is_same<S, R>::value>>
  __host__ __device__ __forceinline__ explicit constexpr Dual(Dual<S> o) : v(R(o.v)), d(R(o.d)) {}

  friend __device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
  friend __device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
  friend __device__ __forceinline__ Dual operator*(Dual a, Dual b) {
    return {a.v * b.v, a.d * b.v + a.v * b.d};
  }
  friend __device__ __forceinline__ Dual operator/(Dual a, Dual b) {
    const R q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
  }
  friend __device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
  __device__ __forceinline__ Dual &operator+=(Dual b) { v += b.v; d += b.d; return *this; }
  __device__ __forceinline__ Dual &operator-=(Dual b) { v -= b.v; d -= b.d; return *this; }
  __device__ __forceinline__ Dual &operator*=(Dual b) { *this = *this * b; return *this; }
  friend __device__ __forceinline__ bool operator>(Dual a, Dual b) { return a.v > b.v; }
  friend __device__ __forceinline__ bool operator<(Dual a, Dual b) { return a.v < b.v; }
  friend __device__ __forceinline__ bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
  friend __device__ __forceinline__ bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
  friend __device__ __forceinline__ bool operator==(Dual a, Dual b) { return a.v == b.v; }
  friend __device__ __forceinline__ bool operator!=(Dual a, Dual b) { return a.v != b.v; }
};

template <typename T> struct is_dual : std::false_type {};
template <typename R> struct is_dual<Dual<R>> : std::true_type {};

// the value of a plain or dual number
__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ float val(float x) { return x; }
template <typename R>
__device__ __forceinline__ R val(Dual<R> x) { return x.v; }

template <typename R>
__device__ __forceinline__ Dual<R> fabs(Dual<R> x) {
  return x.v < R(0) ? Dual<R>{-x.v, -x.d} : x;
}
template <typename R>
__device__ __forceinline__ Dual<R> sqrt(Dual<R> x) {
  const R s = sqrt(x.v);
  return {s, x.d / (R(2) * s)};
}
template <typename R>
__device__ __forceinline__ Dual<R> rexp(Dual<R> x) {
  const R e = rexp(x.v);
  return {e, e * x.d};
}
template <typename R>
__device__ __forceinline__ void rsincos(Dual<R> x, Dual<R> *s, Dual<R> *c) {
  R sv, cv;
  rsincos(x.v, &sv, &cv);
  *s = {sv, cv * x.d};
  *c = {cv, -sv * x.d};
}
template <typename R>
__device__ __forceinline__ Dual<R> rrcp(Dual<R> x) {
  const R r = rrcp(x.v);
  return {r, -r * r * x.d};
}
template <typename R>
__device__ __forceinline__ Dual<R> rdiv(Dual<R> a, Dual<R> b) {
  const R ib = rrcp(b.v), q = a.v * ib;
  return {q, (a.d - q * b.d) * ib};
}
template <typename R>
__device__ __forceinline__ int humlicek_region(Dual<R> a, Dual<R> v) {
  return humlicek_region<R>(a.v, v.v);
}

// Armstrong's Voigt function (FP64 only) with its derivatives by central
// differences in a and v (the function is analytic; h = 1e-6 keeps both the
// truncation and the rounding error near 1e-10 of the value)
template <typename R>
__device__ __forceinline__ Dual<R> voigt_armstrong_t(Dual<R> a, Dual<R> v) {
  const double av = (double) a.v, vv = (double) v.v, h = 1.0e-6;
  const double H = voigt_armstrong(av, vv);
  const double da = (voigt_armstrong(av + h, vv) - voigt_armstrong(av > h ? av - h : av, vv)) /
                    (av > h ? 2.0 * h : h);
  const double dv = (voigt_armstrong(av, vv + h) - voigt_armstrong(av, vv - h)) / (2.0 * h);
  return {(R) H, (R) (da * (double) a.d + dv * (double) v.d)};
}
__device__ __forceinline__ double voigt_armstrong_t(double a, double v) { return voigt_armstrong(a, v); }
__device__ __forceinline__ float voigt_armstrong_t(float a, float v) {
  return (float) voigt_armstrong((double) a, (double) v);
}

// Cubic Bezier weights with their derivatives in dt: the derivative of
// exactly the branch bezier3() evaluates.  Closed forms N(t)/t^3 give
// (N'(t) - 3 N(t)/t) / t^3 with
//   N_alpha' = 6 - 6t + 3t^2 - 6e,     N_beta'  = t^3 e,
//   N_gamma' = 3 (-4 + 2t + (4 + 2t) e),  N_theta' = 3 (2 - (2 + 2t + t^2) e),  e = exp(-t);
// series are differentiated term by term.
template <typename R>
__device__ __forceinline__ Bezier3<Dual<R>> bezier3(Dual<R> dt) {
  const Bezier3<R> b = bezier3(dt.v);
  const R t = dt.v;
  R da, db, dg, dth, de;
  constexpr bool fp32 = std::is_same<R, float>::value;
  if (fp32 ? (t < R(1)) : (t < R(5.e-2))) {
    if constexpr (fp32) {
      float a = 0.0f, be = 0.0f, g = 0.0f, th = 0.0f;
#pragma unroll
      for (int m = BEZ_NSERIES; m >= 1; m--) {
        a = a * t + (float) (m * 6.0 * e_k(m + 3));
        be = be * t + (float) (m * -(6.0 * e_k(m + 3) + 6.0 * e_k(m + 2) + 3.0 * e_k(m + 1) + e_k(m)));
        g = g * t + (float) (m * -3.0 * (6.0 * e_k(m + 3) + 2.0 * e_k(m + 2)));
        th = th * t + (float) (m * 3.0 * (6.0 * e_k(m + 3) + 4.0 * e_k(m + 2) + e_k(m + 1)));
      }
      da = a; db = be; dg = g; dth = th;
      de = -b.eps;
    } else {
      // the reference's 4th-order branch (rh/bezier_aux.c:350-357)
      const R t2 = t * t, t3 = t2 * t;
      da = 0.25 - 0.10 * t + t2 / 40.0 - t3 / 210.0;
      db = 0.25 - 0.40 * t + t2 / 4.0 - t3 * (4.0 / 42.0);
      dg = 0.25 - 0.20 * t + t2 * 0.075 - t3 * (4.0 / 210.0);
      dth = 0.25 - 0.30 * t + t2 * 0.15 - t3 * (4.0 / 84.0);
      de = -1.0 + t - 0.5 * t2 + t3 / 6.0;
    }
  } else {
    const R e = b.eps, it = R(1) / t, it3 = it * it * it;
    da = (R(6) - R(6) * t + R(3) * t * t - R(6) * e) * it3 - R(3) * b.alpha * it;
    db = e - R(3) * b.beta * it;
    dg = R(3) * (R(-4) + R(2) * t + (R(4) + R(2) * t) * e) * it3 - R(3) * b.gamma * it;
    dth = R(3) * (R(2) - (R(2) + R(2) * t + t * t) * e) * it3 - R(3) * b.theta * it;
    de = -e;
  }
  Bezier3<Dual<R>> o;
  o.alpha = {b.alpha, da * dt.d};
  o.beta = {b.beta, db * dt.d};
  o.gamma = {b.gamma, dg * dt.d};
  o.theta = {b.theta, dth * dt.d};
  o.eps = {b.eps, de * dt.d};
  return o;
}

// w3 weights (rh/w3.c:44-63) with derivatives in dtau: w0' = e, w1' = dtau e
// (e = exp(-dtau)); the small-dtau series and the large-dtau limits are
// differentiated as written.
__device__ __forceinline__ void w3d(double dtau, double *w, double *dw) {
  w3(dtau, w);
  if (dtau < 5.0E-4) {
    dw[0] = 1.0 - dtau;
    dw[1] = dtau * (1.0 - dtau);
  } else if (dtau > 50.0) {
    dw[0] = dw[1] = 0.0;
  } else {
    const double e = exp(-dtau);
    dw[0] = e;
    dw[1] = dtau * e;
  }
}


// ---------------------------------------------------------------------------
// Multi-tangent dual numbers: a value and NP directional derivatives, for
// the response functions' chain to the atmosphere parameters (one pass
// carries every parameter's tangent; the value is computed once).
template <typename R, int NP>
struct DualN {
  R v, d[NP];
  DualN() = default;
  template <typename S, typename = std::enable_if_t<std::is_arithmetic<S>::value>>
  __device__ __forceinline__ DualN(S val) : v(R(val)) {
#pragma unroll
    for (int i = 0; i < NP; i++) d[i] = R(0);
  }
  template <typename S, typename = std::enable_if_t<!std::is_same<S, R>::value>>
  __device__ __forceinline__ explicit DualN(const DualN<S, NP> &o) : v(R(o.v)) {
#pragma unroll
    for (int i = 0; i < NP; i++) d[i] = R(o.d[i]);
  }
  friend __device__ __forceinline__ DualN operator+(const DualN &a, const DualN &b) {
    DualN r;
    r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < NP; i++) r.d[i] = a.d[i] + b.d[i];
    return r;
  }
  friend __device__ __forceinline__ DualN operator-(const DualN &a, const DualN &b) {
    DualN r;
    r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < NP; i++) r.d[i] = a.d[i] - b.d[i];
    return r;
  }
  friend __device__ __forceinline__ DualN operator*(const DualN &a, const DualN &b) {
    DualN r;
    r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < NP; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
  }
  friend __device__ __forceinline__ DualN operator/(const DualN &a, const DualN &b) {
    DualN r;
    const R ib = R(1) / b.v;
    r.v = a.v * ib;
#pragma unroll
    for (int i = 0; i < NP; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
  }
  friend __device__ __forceinline__ DualN operator-(const DualN &a) {
    DualN r;
    r.v = -a.v;
#pragma unroll
    for (int i = 0; i < NP; i++) r.d[i] = -a.d[i];
    return r;
  }
  __device__ __forceinline__ DualN &operator+=(const DualN &b) { *this = *this + b; return *this; }
  __device__ __forceinline__ DualN &operator-=(const DualN &b) { *this = *this - b; return *this; }
  __device__ __forceinline__ DualN &operator*=(const DualN &b) { *this = *this * b; return *this; }
  friend __device__ __forceinline__ bool operator>(const DualN &a, const DualN &b) { return a.v > b.v; }
  friend __device__ __forceinline__ bool operator<(const DualN &a, const DualN &b) { return a.v < b.v; }
  friend __device__ __forceinline__ bool operator>=(const DualN &a, const DualN &b) { return a.v >= b.v; }
  friend __device__ __forceinline__ bool operator<=(const DualN &a, const DualN &b) { return a.v <= b.v; }
  friend __device__ __forceinline__ bool operator==(const DualN &a, const DualN &b) { return a.v == b.v; }
  friend __device__ __forceinline__ bool operator!=(const DualN &a, const DualN &b) { return a.v != b.v; }
};

template <typename R, int NP>
__device__ __forceinline__ R val(const DualN<R, NP> &x) { return x.v; }

// f(x) with f(x.v) = fv and f'(x.v) = fd
template <typename R, int NP>
__device__ __forceinline__ DualN<R, NP> dn_apply(const DualN<R, NP> &x, R fv, R fd) {
  DualN<R, NP> r;
  r.v = fv;
#pragma unroll
  for (int i = 0; i < NP; i++) r.d[i] = fd * x.d[i];
  return r;
}
template <typename R, int NP>
__device__ __forceinline__ DualN<R, NP> fabs(const DualN<R, NP> &x) {
  return dn_apply(x, x.v < R(0) ? -x.v : x.v, x.v < R(0) ? R(-1) : R(1));
}
template <typename R, int NP>
__device__ __forceinline__ DualN<R, NP> rexp(const DualN<R, NP> &x) {
  const R e = rexp(x.v);
  return dn_apply(x, e, e);
}
template <typename R, int NP>
__device__ __forceinline__ DualN<R, NP> rrcp(const DualN<R, NP> &x) {
  const R r = rrcp(x.v);
  return dn_apply(x, r, -r * r);
}
template <typename R, int NP>
__device__ __forceinline__ DualN<R, NP> rdiv(const DualN<R, NP> &a, const DualN<R, NP> &b) {
  return a * rrcp(b);
}

// Armstrong's function with its partials in a and v (as for Dual)
template <typename R, int NP>
__device__ __forceinline__ DualN<R, NP> voigt_armstrong_t(const DualN<R, NP> &a,
                                                          const DualN<R, NP> &v) {
  const double av = (double) a.v, vv = (double) v.v, h = 1.0e-6;
  const double H = voigt_armstrong(av, vv);
  const double da = (voigt_armstrong(av + h, vv) - voigt_armstrong(av > h ? av - h : av, vv)) /
                    (av > h ? 2.0 * h : h);
  const double dv = (voigt_armstrong(av, vv + h) - voigt_armstrong(av, vv - h)) / (2.0 * h);
  DualN<R, NP> r;
  r.v = (R) H;
#pragma unroll
  for (int i = 0; i < NP; i++) r.d[i] = (R) (da * (double) a.d[i] + dv * (double) v.d[i]);
  return r;
}

// The Humlicek functions of the line profiles.  A plain or dual argument
// evaluates voigt_humlicek() in that type.  For NP tangents: w = H + iF is
// an analytic function of z = a - iv inside each region (rational, region
// IV with exp(z^2)), so one evaluation on a dual number seeded in a gives
// w' = dw/dz = (H_a, F_a), and dH = H_a da + F_a dv, dF = F_a da - H_a dv
// (Cauchy-Riemann) for every tangent.
// `region` is Humlicek's region of the argument (humlicek_region).
template <typename T>
struct VoigtEval {
  static __device__ __forceinline__ T eval(T a, T v, int region, T *F) {
    return voigt_humlicek_r<T>(a, v, region, F);
  }
};
template <typename R, int NP>
struct VoigtEval<DualN<R, NP>> {
  static __device__ __forceinline__ DualN<R, NP> eval(const DualN<R, NP> &a,
                                                      const DualN<R, NP> &v, int region,
                                                      DualN<R, NP> *F) {
    Dual<R> Fd;
    const Dual<R> Hd =
        voigt_humlicek_r<Dual<R>>(Dual<R>(a.v, R(1)), Dual<R>(v.v, R(0)), region, &Fd);
    const R Ha = Hd.d, Fa = Fd.d;
    DualN<R, NP> H;
    H.v = Hd.v;
    F->v = Fd.v;
#pragma unroll
    for (int i = 0; i < NP; i++) {
      H.d[i] = Ha * a.d[i] + Fa * v.d[i];
      F->d[i] = Fa * a.d[i] - Ha * v.d[i];
    }
    return H;
  }
};

}  // namespace rhg
