// Device-side numerics shared by the gfx950 kernels.
//
// The FP64 ("plain") path evaluates every expression in the order of the
// reference routine it restates (cited per function) and is compiled with
// FP contraction disabled, so that the only differences from the CPU are
// the last-bit differences of the GPU's exp/log/pow.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pyrh_amd.h"

namespace rhg {

// rh/constant.h:28-73, rh/rh.h:36-50
constexpr double CLIGHT = 2.99792458E+08;
constexpr double HPLANCK = 6.6260755E-34;
constexpr double KBOLTZMANN = 1.380658E-23;
constexpr double AMU = 1.6605402E-27;
constexpr double M_ELECTRON = 9.1093897E-31;
constexpr double Q_ELECTRON = 1.60217733E-19;
constexpr double EPSILON_0 = 8.854187817E-12;
constexpr double EV = 1.60217733E-19;
constexpr double THETA0 = 5.03974756E+03;
constexpr double NM_TO_M = 1.0E-09;
constexpr double CM_TO_M = 1.0E-02;
constexpr double KM_TO_M = 1.0E+03;
constexpr double G_TO_KG = 1.0E-03;
constexpr double ERG_TO_JOULE = 1.0E-07;
constexpr double PI = 3.14159265358979;
constexpr double SQRTPI = 1.77245385090551;
constexpr double INV_SQRTPI = 1.0 / SQRTPI;
constexpr double LG10 = 2.30258509299404568402;
constexpr double MAX_EXPONENT = 150.0;
constexpr double MAX_GAUSS_DOPPLER = 7.0;

constexpr int N_MAX_CHEM_ITER = 10;
constexpr double CHEM_ITER_LIMIT = 1.0E-3;
constexpr int NSTAGE_MAX = 8;
constexpr int NNUC_MAX = 8;     // nuclei bound in molecules
constexpr int NMOL_MAX = 24;

constexpr int NTHETA = 16, NFF = 17, NTHETA_H2 = 8, NFF_H2 = 19;
constexpr int NTEMP_H2P = 10, NFF_H2P = 15, NTOH = 15, NEOH = 130, NTCH = 15, NECH = 105;

__device__ __forceinline__ double pow10d(double x) { return exp(LG10 * x); }

// rh/hunt.c:92-117, ascending table
__device__ __forceinline__ int locate_asc(int n, const double *a, double v) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    int idx = (hi + lo) >> 1;
    if (v >= a[idx]) lo = idx; else hi = idx;
  }
  return lo;
}

// rh/linear.c:22-54 for one abscissa (ascending table)
__device__ __forceinline__ double linear1(int n, const double *xt, const double *yt, double x) {
  if (x <= xt[0]) return yt[0];
  if (x >= xt[n - 1]) return yt[n - 1];
  int j = locate_asc(n, xt, x);
  double fx = (xt[j + 1] - x) / (xt[j + 1] - xt[j]);
  return fx * yt[j] + (1 - fx) * yt[j + 1];
}

// linear1 at one abscissa x for several tables yt on the same xt: the
// bracket and weight are found once, with linear1's arithmetic, so each
// value is bit-identical to linear1(n, xt, yt, x).
struct Linear1At {
  int n, j, edge;   // edge -1: x <= xt[0]; 1: x >= xt[n - 1]; 0: inside
  double fx;
  __device__ __forceinline__ Linear1At(int n_, const double *xt, double x)
      : n(n_), j(0), edge(0), fx(0.0) {
    if (x <= xt[0]) {
      edge = -1;
    } else if (x >= xt[n - 1]) {
      edge = 1;
    } else {
      j = locate_asc(n, xt, x);
      fx = (xt[j + 1] - x) / (xt[j + 1] - xt[j]);
    }
  }
  __device__ __forceinline__ double operator()(const double *yt) const {
    if (edge < 0) return yt[0];
    if (edge > 0) return yt[n - 1];
    return fx * yt[j] + (1 - fx) * yt[j + 1];
  }
};

// rh/hydrogen.c:998-1020
__device__ __forceinline__ double bilinear(int Ncol, int Nrow, const double *f, double x, double y) {
  int i = (int) x;
  double fx = x - i;
  int i1 = (i == Ncol - 1) ? i : i + 1;
  int j = (int) y;
  double fy = y - j;
  int j1 = (j == Nrow - 1) ? j : j + 1;
  return (1.0 - fx) * (1.0 - fy) * f[j * Ncol + i] + fx * (1.0 - fy) * f[j * Ncol + i1] +
         (1.0 - fx) * fy * f[j1 * Ncol + i] + fx * fy * f[j1 * Ncol + i1];
}

// clamped fractional index (rh/hydrogen.c:509-523)
__device__ __forceinline__ double clamp_index(int n, const double *t, double v) {
  if (v <= t[0]) return 0;
  if (v >= t[n - 1]) return n - 1;
  int i = locate_asc(n, t, v);
  return (double) i + (v - t[i]) / (t[i + 1] - t[i]);
}

// ---------------------------------------------------------------------------
// Voigt and Faraday-Voigt functions: Humlicek (1982) regions I-IV with the
// region test of rh/voigt.c:396-413 and the rational forms of
// rh/humlicek.c:28-120, in complex arithmetic written out in real pairs.
// R = double follows the products and the quotient of rh/complex.c:45-100;
// R = float is the FP32 form.
template <typename T>
struct cpx { T r, i; };

template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}

// Reciprocal and quotient: IEEE-exact in FP64; in the FP32 form the
// hardware reciprocal (v_rcp_f32, 1 ulp), which the form's 1e-4 tolerance
// allows.
//
// A translation unit of the FP32 form defines RHG_FP32_FORM_TU before
// including this header: its FP64 arithmetic (the recursion and the points,
// which need FP64 range and precision but not the reference's rounding)
// takes the reciprocal as the hardware estimate refined by two Newton steps
// (relative error ~1e-16, not correctly rounded) instead of the IEEE
// division sequence, and the Bezier weights below from their series.
#ifdef RHG_FP32_FORM_TU
__device__ __forceinline__ double rrcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ double rdiv(double a, double b) { return a * rrcp(b); }
constexpr bool kFp32FormTU = true;
#else
__device__ __forceinline__ double rrcp(double x) { return 1.0 / x; }
__device__ __forceinline__ double rdiv(double a, double b) { return a / b; }
constexpr bool kFp32FormTU = false;
#endif
__device__ __forceinline__ float rrcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// a / b with one reciprocal of |b|^2
template <typename T>
__device__ __forceinline__ cpx<T> cdiv(cpx<T> a, cpx<T> b) {
  const T id = rrcp(b.r * b.r + b.i * b.i);
  return {(a.r * b.r + a.i * b.i) * id, (a.i * b.r - a.r * b.i) * id};
}

__device__ __forceinline__ double rexp(double x) { return exp(x); }
__device__ __forceinline__ float rexp(float x) { return __expf(x); }
__device__ __forceinline__ void rsincos(double x, double *s, double *c) { sincos(x, s, c); }
// FP32 form: the hardware sine/cosine (the argument here is 2 a v with
// a v small inside Humlicek region IV)
__device__ __forceinline__ void rsincos(float x, float *s, float *c) {
  *s = __sinf(x);
  *c = __cosf(x);
}

// Humlicek's region (0..3 for I..IV) of the point z = a - i v, the test of
// rh/voigt.c:396-413; voigt_humlicek() branches on exactly this test.
// The three comparisons are all evaluated and combined with selects, so
// that a wavefront does not branch to find each lane's region.
template <typename R>
__device__ __forceinline__ int humlicek_region(R a, R v) {
  const R s = fabs(v) + a;
  const int inner = (a >= R(0.195) * fabs(v) - R(0.176)) ? 2 : 3;
  const int mid = (s >= R(5.5)) ? 1 : inner;
  return (s >= R(15.0)) ? 0 : mid;
}

// Each region ends in one complex quotient num/den (region IV: z num/den,
// subtracted from exp(-u)); the branches form num and den only and the one
// quotient is taken after they rejoin, so a wavefront whose lanes sit in
// several regions pays for the quotient once.  The operations and their
// order per lane are those of the per-region formulas.
// voigt_humlicek_r takes the region from the caller (the FP32 form tests
// the region on the FP64 argument, so that both forms evaluate the same
// approximation).
template <typename R>
__device__ __forceinline__ R voigt_humlicek_r(R a, R v, int region, R *F) {
  cpx<R> z = {a, -v}, z1, z2, u, e = {R(0), R(0)};
  if (region == 0) {
    z1 = {R(0.5641896) * z.r, R(0.5641896) * z.i};
    z2 = cmul(z, z);
    z2.r = z2.r + R(0.5);
  } else if (region == 1) {
    u = cmul(z, z);
    z1 = {R(0.5641896) * u.r, R(0.5641896) * u.i};
    z1.r = z1.r + R(1.410474);
    z1 = cmul(z, z1);
    cpx<R> u3 = {u.r + R(3.0), u.i};
    z2 = cmul(u, u3);
    z2.r = z2.r + R(0.75);
  } else if (region == 2) {
    const R a3[5] = {R(0.5642236), R(3.778987), R(11.96482), R(20.20933), R(16.4955)};
    const R b3[5] = {R(6.699398), R(21.69274), R(39.27121), R(38.82363), R(16.4955)};
    z1 = {a3[0], R(0.0)};
    z2 = {z.r + b3[0], z.i};
#pragma unroll
    for (int n = 1; n < 5; n++) {
      z1 = cmul(z1, z); z1.r = z1.r + a3[n];
      z2 = cmul(z2, z); z2.r = z2.r + b3[n];
    }
  } else {
    const R a4[7] = {R(0.56419), R(1.320522), R(35.7668), R(219.031), R(1540.787), R(3321.99),
                     R(36183.31)};
    const R b4[7] = {R(1.841439), R(61.57037), R(364.2191), R(2186.181), R(9022.228), R(24322.84),
                     R(32066.6)};
    cpx<R> mz = {R(-1.0) * z.r, R(-1.0) * z.i};
    u = cmul(z, mz);
    z1 = {a4[0], R(0.0)};
    z2 = {u.r + b4[0], u.i};
#pragma unroll
    for (int n = 1; n < 7; n++) {
      z1 = cmul(u, z1); z1.r = z1.r + a4[n];
      z2 = cmul(u, z2); z2.r = z2.r + b4[n];
    }
    cpx<R> mu = {R(-1.0) * u.r, R(-1.0) * u.i};
    R ex = rexp(mu.r), sn, cs;
    rsincos(mu.i, &sn, &cs);
    e = {ex * cs, ex * sn};
    z1 = cmul(z, z1);
  }
  cpx<R> W = cdiv(z1, z2);
  if (region == 3) W = {e.r - W.r, e.i - W.i};
  *F = W.i;
  return W.r;
}

template <typename R>
__device__ __forceinline__ R voigt_humlicek(R a, R v, R *F) {
  return voigt_humlicek_r<R>(a, v, humlicek_region(a, v), F);
}

// ---------------------------------------------------------------------------
// Armstrong Voigt function, rh/voigt.c:126-243
__device__ constexpr double c_arm_t[10] = {0.2453407083, 0.7374737285, 1.2340762153, 1.7385377121,
                                   2.2549740020, 2.7888060584, 3.3478545673, 3.9447640401,
                                   4.6036824495, 5.3874808900};
__device__ constexpr double c_arm_w[10] = {4.6224366960e-01, 2.8667550536e-01, 1.0901720602e-01,
                                   2.4810520887e-02, 3.2437733422e-03, 2.2833863601e-04,
                                   7.8025564785e-06, 1.0860693707e-07, 4.3993409922e-10,
                                   2.2293936455e-13};
__device__ constexpr double c_k1[34] = {
    0.1999999999972224, -0.1840000000029998, 0.1558399999965025, -0.1216640000043988,
    0.0877081599940391, -0.0585141248086907, 0.0362157301623914, -0.0208497654398036,
    0.0111960116346270, -0.56231896167109e-02, 0.26487634172265e-02, -0.11732670757704e-02,
    0.4899519978088e-03, -0.1933630801528e-03, 0.722877446788e-04, -0.256555124979e-04,
    0.86620736841e-05, -0.27876379719e-05, 0.8566873627e-06, -0.2518433784e-06,
    0.709360221e-07, -0.191732257e-07, 0.49801256e-08, -0.12447734e-08, 0.2997777e-09,
    -0.696450e-10, 0.156262e-10, -0.33897e-11, 0.7116e-12, -0.1447e-12, 0.285e-13,
    -0.55e-14, 0.10e-14, -0.2e-15};

__device__ __noinline__ double voigt_armstrong(double a, double v) {
  if (v < 0.0) v = -v;
  if ((a < 1.0 && v < 4.0) || (a < 1.8 / (v + 1.0))) {
    const double TWOSQRTPI = 1.12837917, EXPMAX = 70.0, TINY = 1.0E-08;
    double a2 = a * a, v2 = v * v, u1, dn01, dn02, dn, funct, bn = 0.0;
    if ((v2 - a2) > EXPMAX) u1 = 0.0;
    else u1 = exp(a2 - v2) * cos(2.0 * v * a);
    if (v > 5.0) {
      double v2i = 1.0 / v2;
      dn01 = -v2i * (0.5 + v2i * (0.75 + v2i * (1.875 + v2i * (6.5625 +
             v2i * (29.53125 + v2i * (1162.4218 + v2i * 1055.7421))))));
      dn02 = (1.0 - dn01) / (2.0 * v);
    } else {
      double bn01 = 0.0, bn02 = 0.0, v1 = v / 5.0, coef = 4.0 * v1 * v1 - 2.0;
      for (int n = 33; n >= 0; n--) {
        bn = coef * bn01 - bn02 + c_k1[n];
        bn02 = bn01;
        bn01 = bn;
      }
      dn02 = v1 * (bn - bn02);
      dn01 = 1.0 - 2.0 * v * dn02;
    }
    funct = a * dn01;
    if (a > TINY) {
      double q = 1.0, an = a;
      for (int n = 2; n <= 50; n++) {
        dn = (v * dn01 + dn02) * (-2.0 / n);
        dn02 = dn01;
        dn01 = dn;
        if (n % 2) {
          q = -q;
          an *= a2;
          double g = dn * an;
          funct += q * g;
          if (fabs(g / funct) <= TINY) return (u1 - TWOSQRTPI * funct);
        }
      }
    }
    return (u1 - TWOSQRTPI * funct);
  } else if (a < 2.5 && v < 4.0) {
    double g = 0.0, a2 = a * a;
    for (int n = 0; n < 10; n++) {
      double r = c_arm_t[n] - v, s = c_arm_t[n] + v;
      g += (4.0 * c_arm_t[n] * c_arm_t[n] - 2.0) *
           (r * atan(r / a) + s * atan(s / a) - 0.5 * a * (log(a2 + r * r) + log(a2 + s * s))) *
           c_arm_w[n];
    }
    return g / PI;
  } else {
    double g = 0.0, a2 = a * a;
    for (int n = 0; n < 10; n++) {
      double d1 = v - c_arm_t[n], d2 = v + c_arm_t[n];
      g += (1.0 / (d1 * d1 + a2) + 1.0 / (d2 * d2 + a2)) * c_arm_w[n];
    }
    return (a * g) / PI;
  }
}


// ---------------------------------------------------------------------------
// rh/planck.c:38-63
__device__ __forceinline__ double planck(double T, double hc_kla, double twohnu3_c2) {
  double hc_Tkla = hc_kla / T;
  return (hc_Tkla <= MAX_EXPONENT) ? twohnu3_c2 / (exp(hc_Tkla) - 1.0) : 0.0;
}

// rh/w3.c:44-63
__device__ __forceinline__ void w3(double dtau, double *w) {
  if (dtau < 5.0E-4) {
    w[0] = dtau * (1.0 - 0.5 * dtau);
    double delta = dtau * dtau;
    w[1] = delta * (0.5 - dtau / 3.0);
    delta *= dtau;
    w[2] = delta * (1.0 / 3.0 - 0.25 * dtau);
  } else if (dtau > 50.0) {
    w[1] = w[0] = 1.0;
    w[2] = 2.0;
  } else {
    double expdt = exp(-dtau);
    w[0] = 1.0 - expdt;
    w[1] = w[0] - dtau * expdt;
    w[2] = 2.0 * w[1] - dtau * dtau * expdt;
  }
}

// Fritsch & Butland (1984) monotone central derivative, rh/bezier_aux.c:34-51.
// CentDeriv holds what depends only on the two intervals, so that the
// derivatives of several quantities on the same intervals share it.
template <typename R>
struct CentDeriv {
  R idup, idn, alpha;
  __device__ __forceinline__ CentDeriv(R dsup, R dsdn)
      : idup(rrcp(dsup)), idn(rrcp(dsdn)),
        alpha(R(0.333333333333333333333333) * (R(1.0) + rdiv(dsdn, dsdn + dsup))) {}
  __device__ __forceinline__ R operator()(R up, R c, R dn) const {
    const R fim1 = (c - up) * idup, fi = (dn - c) * idn;
    if (fim1 * fi > R(0)) return rdiv(fim1 * fi, (R(1.0) - alpha) * fim1 + alpha * fi);
    return R(0.0);
  }
};

template <typename R>
__device__ __forceinline__ R cent_deriv(R dsup, R dsdn, R up, R c, R dn) {
  return CentDeriv<R>(dsup, dsdn)(up, c, dn);
}

// cubic Bezier integration weights, rh/bezier_aux.c:333-359
__device__ __forceinline__ void bezier3_coeffs(double dt, double &alpha, double &beta,
                                               double &gamma, double &theta, double &eps) {
  double dt2 = dt * dt, dt3 = dt2 * dt;
  if (dt >= 5.e-2) {
    eps = exp(-dt);
    alpha = (-6.0 + 6.0 * dt - 3.0 * dt2 + dt3 + 6.0 * eps) / dt3;
    dt3 = 1.0 / dt3;
    beta = (6.0 + (-6.0 - dt * (6.0 + dt * (3.0 + dt))) * eps) * dt3;
    gamma = 3.0 * (6.0 + (-4.0 + dt) * dt - 2.0 * (3.0 + dt) * eps) * dt3;
    theta = 3.0 * (eps * (6.0 + dt2 + 4.0 * dt) + 2.0 * dt - 6.0) * dt3;
  } else {
    double dt4 = dt2 * dt2;
    eps = 1.0 - dt + 0.5 * dt2 - dt3 / 6.0 + dt4 / 24.0;
    alpha = 0.25 * dt - 0.05 * dt2 + dt3 / 120.0 - dt4 / 840.0;
    beta = 0.25 * dt - 0.20 * dt2 + dt3 / 12.0 - dt4 / 42.0;
    gamma = 0.25 * dt - 0.10 * dt2 + dt3 * 0.025 - dt4 / 210.0;
    theta = 0.25 * dt - 0.15 * dt2 + dt3 * 0.05 - dt4 / 84.0;
  }
}

// Cubic Bezier weights (alpha, beta, gamma, theta, eps = exp(-dt)) in the
// precision of the recursion.  FP64: the reference's branches,
// bezier3_coeffs().  FP32: the closed forms cancel (at dt = 0.05 the
// numerator of alpha is 1.5e-6 of its terms), so below dt = 1 the weights
// are their Taylor series about 0, whose coefficients follow from
// exp(-t) = sum e_k t^k with e_k = (-1)^k / k!:
//   alpha_m = 6 e_{m+3},  beta_m = -(6 e_{m+3} + 6 e_{m+2} + 3 e_{m+1} + e_m),
//   gamma_m = -3 (6 e_{m+3} + 2 e_{m+2}),  theta_m = 3 (6 e_{m+3} + 4 e_{m+2} + e_{m+1})
// (m = 1..10; the first four terms are the reference's 4th-order branch),
// and from dt = 1 on the closed forms, whose cancellation is then < 30x.
template <typename R>
struct Bezier3 {
  R alpha, beta, gamma, theta, eps;
};

constexpr double e_k(int k) {
  double f = 1.0;
  for (int i = 2; i <= k; i++) f *= i;
  return ((k & 1) ? -1.0 : 1.0) / f;
}
constexpr int BEZ_NSERIES = 10;

// FP64 weights: the reference's branches, bezier3_coeffs().  In a
// translation unit of the FP32 form the same branches with the quotients by
// constants taken as products and 1/dt^3 by rrcp (no IEEE division).
__device__ __forceinline__ Bezier3<double> bezier3(double dt) {
  Bezier3<double> b;
  if constexpr (kFp32FormTU) {
    const double dt2 = dt * dt, dt3 = dt2 * dt;
    if (dt >= 5.e-2) {
      const double e = exp(-dt), idt3 = rrcp(dt3);
      b.eps = e;
      b.alpha = (-6.0 + 6.0 * dt - 3.0 * dt2 + dt3 + 6.0 * e) * idt3;
      b.beta = (6.0 + (-6.0 - dt * (6.0 + dt * (3.0 + dt))) * e) * idt3;
      b.gamma = 3.0 * (6.0 + (-4.0 + dt) * dt - 2.0 * (3.0 + dt) * e) * idt3;
      b.theta = 3.0 * (e * (6.0 + dt2 + 4.0 * dt) + 2.0 * dt - 6.0) * idt3;
    } else {
      const double dt4 = dt2 * dt2;
      b.eps = 1.0 - dt + 0.5 * dt2 - dt3 * (1.0 / 6.0) + dt4 * (1.0 / 24.0);
      b.alpha = 0.25 * dt - 0.05 * dt2 + dt3 * (1.0 / 120.0) - dt4 * (1.0 / 840.0);
      b.beta = 0.25 * dt - 0.20 * dt2 + dt3 * (1.0 / 12.0) - dt4 * (1.0 / 42.0);
      b.gamma = 0.25 * dt - 0.10 * dt2 + dt3 * 0.025 - dt4 * (1.0 / 210.0);
      b.theta = 0.25 * dt - 0.15 * dt2 + dt3 * 0.05 - dt4 * (1.0 / 84.0);
    }
  } else {
    bezier3_coeffs(dt, b.alpha, b.beta, b.gamma, b.theta, b.eps);
  }
  return b;
}

__device__ __forceinline__ Bezier3<float> bezier3(float dt) {
  Bezier3<float> b;
  b.eps = __expf(-dt);
  if (dt < 1.0f) {
    float a = 0.0f, be = 0.0f, g = 0.0f, th = 0.0f;
#pragma unroll
    for (int m = BEZ_NSERIES; m >= 1; m--) {
      a = a * dt + (float) (6.0 * e_k(m + 3));
      be = be * dt + (float) (-(6.0 * e_k(m + 3) + 6.0 * e_k(m + 2) + 3.0 * e_k(m + 1) + e_k(m)));
      g = g * dt + (float) (-3.0 * (6.0 * e_k(m + 3) + 2.0 * e_k(m + 2)));
      th = th * dt + (float) (3.0 * (6.0 * e_k(m + 3) + 4.0 * e_k(m + 2) + e_k(m + 1)));
    }
    b.alpha = a * dt;
    b.beta = be * dt;
    b.gamma = g * dt;
    b.theta = th * dt;
  } else {
    const float dt2 = dt * dt, idt3 = rrcp(dt2 * dt), e = b.eps;
    b.alpha = (-6.0f + 6.0f * dt - 3.0f * dt2 + dt2 * dt + 6.0f * e) * idt3;
    b.beta = (6.0f + (-6.0f - dt * (6.0f + dt * (3.0f + dt))) * e) * idt3;
    b.gamma = 3.0f * (6.0f + (-4.0f + dt) * dt - 2.0f * (3.0f + dt) * e) * idt3;
    b.theta = 3.0f * (e * (6.0f + dt2 + 4.0f * dt) + 2.0f * dt - 6.0f) * idt3;
  }
  return b;
}

}  // namespace rhg
