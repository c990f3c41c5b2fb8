/* Test oracle: CPU restatement of RH's LTE synthesis as pyrh runs it.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C and FP64, the
 * reference algorithm of rhf1d() with no active atom
 * (rh/rhf1d/pyrh_compute1dray.c:112-389) for one column at a time, reading
 * the host tables of rhg_model_t.  It is used as the checker of the GPU
 * library and as the CPU baseline; the product never calls it.
 *
 * Floating-point operations are written in the order the reference
 * evaluates them (operator precedence of its macros included), and libm is
 * called for the same functions, so that on the same inputs results agree
 * with the reference to round-off.  Every function cites the reference
 * source it follows.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "rh_oracle.h"

/* rh/constant.h:28-73, rh/rh.h:36-50 */
#define CLIGHT      2.99792458E+08
#define HPLANCK     6.6260755E-34
#define KBOLTZMANN  1.380658E-23
#define AMU         1.6605402E-27
#define M_ELECTRON  9.1093897E-31
#define Q_ELECTRON  1.60217733E-19
#define EPSILON_0   8.854187817E-12
#define EV          1.60217733E-19
#define THETA0      5.03974756E+03
#define NM_TO_M     1.0E-09
#define CM_TO_M     1.0E-02
#define KM_TO_M     1.0E+03
#define G_TO_KG     1.0E-03
#define ERG_TO_JOULE 1.0E-07
#define PI          3.14159265358979
#define SQRTPI      1.77245385090551
#define LG10        2.30258509299404568402
#define RBOHR       5.29177349E-11
#define E_RYDBERG   2.1798741E-18
#define SQ(x)       ((x)*(x))
#define CUBE(x)     ((x)*(x)*(x))
#define POW10(x)    exp(LG10 * (x))
#define MAXV(a, b)  (((a) > (b)) ? (a) : (b))

#define N_MAX_CHEM_ITER 10        /* rh/background.h:18-19 */
#define CHEM_ITER_LIMIT 1.0E-3
#define MAX_EXPONENT    150.0     /* rh/planck.c:29 */
#define MAX_GAUSS_DOPPLER 7.0     /* rh/kurucz.c:92 */

#define NTHETA 16
#define NFF 17
#define NTHETA_H2 8
#define NFF_H2 19
#define NTEMP_H2P 10
#define NFF_H2P 15
#define NTOH 15
#define NEOH 130
#define NTCH 15
#define NECH 105

/* ------------------------------------------------------------------ */
/* numerical utilities                                                 */

/* rh/hunt.c:92-117 (ascending branch; all tables here ascend) */
static int locate_asc(int n, const double *a, double v)
{
  int lo = 0, hi = n, idx;
  if (a[n - 1] > a[0]) {
    while (hi - lo > 1) {
      idx = (hi + lo) >> 1;
      if (v >= a[idx]) lo = idx; else hi = idx;
    }
  } else {
    while (hi - lo > 1) {
      idx = (hi + lo) >> 1;
      if (v <= a[idx]) lo = idx; else hi = idx;
    }
  }
  return lo;
}

/* rh/linear.c:22-54 for one abscissa */
static double linear1(int n, const double *xt, const double *yt, double x)
{
  int asc = xt[1] > xt[0], j;
  double xmin = asc ? xt[0] : xt[n - 1], xmax = asc ? xt[n - 1] : xt[0], fx;
  if (x <= xmin) return asc ? yt[0] : yt[n - 1];
  if (x >= xmax) return asc ? yt[n - 1] : yt[0];
  j = locate_asc(n, xt, x);
  fx = (xt[j + 1] - x) / (xt[j + 1] - xt[j]);
  return fx * yt[j] + (1 - fx) * yt[j + 1];
}

/* rh/hydrogen.c:998-1020 */
static double bilinear(int Ncol, int Nrow, const double *f, double x, double y)
{
  int i, j, i1, j1;
  double fx, fy;
  i = (int) x;  fx = x - i;
  i1 = (i == Ncol - 1) ? i : i + 1;
  j = (int) y;  fy = y - j;
  j1 = (j == Nrow - 1) ? j : j + 1;
  return (1.0 - fx) * (1.0 - fy) * f[j * Ncol + i] +
         fx * (1.0 - fy) * f[j * Ncol + i1] +
         (1.0 - fx) * fy * f[j1 * Ncol + i] +
         fx * fy * f[j1 * Ncol + i1];
}

/* fractional index into an ascending table with end clamping, as the ff
   routines form it (rh/hydrogen.c:509-523) */
static double clamp_index(int n, const double *t, double v)
{
  int i;
  if (v <= t[0]) return 0;
  if (v >= t[n - 1]) return n - 1;
  i = locate_asc(n, t, v);
  return (double) i + (v - t[i]) / (t[i + 1] - t[i]);
}

/* ------------------------------------------------------------------ */
/* Voigt functions                                                     */

typedef struct { double r, i; } cplx;
static cplx cmul(cplx a, cplx b) { cplx c = {a.r*b.r - a.i*b.i, a.r*b.i + a.i*b.r}; return c; }
static cplx csc(double a, cplx b) { cplx c = {a * b.r, a * b.i}; return c; }
static cplx cadr(cplx a, double b) { cplx c = {a.r + b, a.i}; return c; }
static cplx csub(cplx a, cplx b) { cplx c = {a.r - b.r, a.i - b.i}; return c; }
static cplx cdivi(cplx a, cplx b)   /* rh/complex.c:87-100 */
{
  double d = b.r * b.r + b.i * b.i;
  cplx c = {(a.r * b.r + a.i * b.i) / d, (a.i * b.r - a.r * b.i) / d};
  return c;
}

/* Humlicek (1982) rational approximations, regions I-IV, with the region
   test of rh/voigt.c:396-413 and the coefficients of rh/humlicek.c:28-120 */
double rho_voigt_humlicek(double a, double v, double *F)
{
  static const double a3[5] = {0.5642236, 3.778987, 11.96482, 20.20933, 16.4955};
  static const double b3[5] = {6.699398, 21.69274, 39.27121, 38.82363, 16.4955};
  static const double a4[7] = {0.56419, 1.320522, 35.7668, 219.031, 1540.787,
                               3321.99, 36183.31};
  static const double b4[7] = {1.841439, 61.57037, 364.2191, 2186.181, 9022.228,
                               24322.84, 32066.6};
  cplx z = {a, -v}, W, z1, z2, u;
  double s = fabs(v) + a;
  int n;
  if (s >= 15.0) {
    z1 = csc(0.5641896, z);
    z2 = cadr(cmul(z, z), 0.5);
    W = cdivi(z1, z2);
  } else if (s >= 5.5) {
    u = cmul(z, z);
    z1 = csc(0.5641896, u);
    z1 = cmul(z, cadr(z1, 1.410474));
    z2 = cmul(u, cadr(u, 3.0));
    z2 = cadr(z2, 0.75);
    W = cdivi(z1, z2);
  } else if (a >= 0.195 * fabs(v) - 0.176) {
    z1.r = a3[0]; z1.i = 0.0;
    z2 = cadr(z, b3[0]);
    for (n = 1; n < 5; n++) {
      z1 = cadr(cmul(z1, z), a3[n]);
      z2 = cadr(cmul(z2, z), b3[n]);
    }
    W = cdivi(z1, z2);
  } else {
    cplx e;
    u = cmul(z, csc(-1.0, z));
    z1.r = a4[0]; z1.i = 0.0;
    z2 = cadr(u, b4[0]);
    for (n = 1; n < 7; n++) {
      z1 = cadr(cmul(u, z1), a4[n]);
      z2 = cadr(cmul(u, z2), b4[n]);
    }
    {
      cplx mu = csc(-1.0, u), cs = {cos(mu.i), sin(mu.i)};
      e = csc(exp(mu.r), cs);
    }
    W = csub(e, cdivi(cmul(z, z1), z2));
  }
  if (F) *F = W.i;
  return W.r;
}

/* Armstrong (1967) Voigt function, rh/voigt.c:126-243 */
static const double arm_t[10] = {0.2453407083, 0.7374737285, 1.2340762153,
  1.7385377121, 2.2549740020, 2.7888060584, 3.3478545673, 3.9447640401,
  4.6036824495, 5.3874808900};
static const double arm_w[10] = {4.6224366960e-01, 2.8667550536e-01,
  1.0901720602e-01, 2.4810520887e-02, 3.2437733422e-03, 2.2833863601e-04,
  7.8025564785e-06, 1.0860693707e-07, 4.3993409922e-10, 2.2293936455e-13};

static double voigt_k1(double a, double v)
{
  static const double c[34] = {0.1999999999972224, -0.1840000000029998,
    0.1558399999965025, -0.1216640000043988, 0.0877081599940391,
    -0.0585141248086907, 0.0362157301623914, -0.0208497654398036,
    0.0111960116346270, -0.56231896167109e-02, 0.26487634172265e-02,
    -0.11732670757704e-02, 0.4899519978088e-03, -0.1933630801528e-03,
    0.722877446788e-04, -0.256555124979e-04, 0.86620736841e-05,
    -0.27876379719e-05, 0.8566873627e-06, -0.2518433784e-06, 0.709360221e-07,
    -0.191732257e-07, 0.49801256e-08, -0.12447734e-08, 0.2997777e-09,
    -0.696450e-10, 0.156262e-10, -0.33897e-11, 0.7116e-12, -0.1447e-12,
    0.285e-13, -0.55e-14, 0.10e-14, -0.2e-15};
  const double TWOSQRTPI = 1.12837917, EXPMAX = 70.0, TINY = 1.0E-08;
  double a2 = a * a, v2 = v * v, u1, dn01, dn02, dn, v2i, funct, an, q, g,
         coef, bn01, bn02, bn = 0.0, v1;
  int n;
  if ((v2 - a2) > EXPMAX) u1 = 0.0;
  else u1 = exp(a2 - v2) * cos(2.0 * v * a);
  if (v > 5.0) {
    v2i = 1.0 / v2;
    dn01 = -v2i * (0.5 + v2i * (0.75 + v2i * (1.875 + v2i * (6.5625 +
           v2i * (29.53125 + v2i * (1162.4218 + v2i * 1055.7421))))));
    dn02 = (1.0 - dn01) / (2.0 * v);
  } else {
    bn01 = bn02 = 0.0;
    v1 = v / 5.0;
    coef = 4.0 * v1 * v1 - 2.0;
    for (n = 33; n >= 0; n--) {
      bn = coef * bn01 - bn02 + c[n];
      bn02 = bn01;
      bn01 = bn;
    }
    dn02 = (double) (v1 * (bn - bn02));
    dn01 = 1.0 - 2.0 * v * dn02;
  }
  funct = a * dn01;
  if (a > TINY) {
    q = 1.0;
    an = a;
    for (n = 2; n <= 50; n++) {
      dn = (v * dn01 + dn02) * (-2.0 / n);
      dn02 = dn01;
      dn01 = dn;
      if (n % 2) {
        q = -q;
        an *= a2;
        g = dn * an;
        funct += q * g;
        if (fabs(g / funct) <= TINY) return (u1 - TWOSQRTPI * funct);
      }
    }
  }
  return (u1 - TWOSQRTPI * funct);
}

static double voigt_k2(double a, double v)
{
  double g = 0.0, r, s, a2 = a * a;
  int n;
  for (n = 0; n < 10; n++) {
    r = arm_t[n] - v;
    s = arm_t[n] + v;
    g += (4.0 * arm_t[n] * arm_t[n] - 2.0) * (r * atan(r / a) + s * atan(s / a) -
         0.5 * a * (log(a2 + r * r) + log(a2 + s * s))) * arm_w[n];
  }
  return g / PI;
}

static double voigt_k3(double a, double v)
{
  double g = 0.0, a2 = a * a;
  int n;
  for (n = 0; n < 10; n++)
    g += (1.0 / (SQ(v - arm_t[n]) + a2) + 1.0 / (SQ(v + arm_t[n]) + a2)) * arm_w[n];
  return (a * g) / PI;
}

double rho_voigt_armstrong(double a, double v)
{
  if (v < 0.0) v = -v;
  if ((a < 1.0 && v < 4.0) || (a < 1.8 / (v + 1.0))) return voigt_k1(a, v);
  else if (a < 2.5 && v < 4.0) return voigt_k2(a, v);
  else return voigt_k3(a, v);
}

/* ------------------------------------------------------------------ */
/* formal-solution helpers                                            */

double rho_planck(double T, double lambda)   /* rh/planck.c:38-63 */
{
  double hc_kla = (HPLANCK * CLIGHT) / (KBOLTZMANN * NM_TO_M * lambda);
  double twohnu3_c2 = (2.0 * HPLANCK * CLIGHT) / CUBE(NM_TO_M * lambda);
  double hc_Tkla = hc_kla / T;
  return (hc_Tkla <= MAX_EXPONENT) ? twohnu3_c2 / (exp(hc_Tkla) - 1.0) : 0.0;
}

void rho_w3(double dtau, double *w)   /* rh/w3.c:44-63 */
{
  double expdt, delta;
  if (dtau < 5.0E-4) {
    w[0] = dtau * (1.0 - 0.5 * dtau);
    delta = SQ(dtau);
    w[1] = delta * (0.5 - dtau / 3.0);
    delta *= dtau;
    w[2] = delta * (1.0 / 3.0 - 0.25 * dtau);
  } else if (dtau > 50.0) {
    w[1] = w[0] = 1.0;
    w[2] = 2.0;
  } else {
    expdt = exp(-dtau);
    w[0] = 1.0 - expdt;
    w[1] = w[0] - dtau * expdt;
    w[2] = 2.0 * w[1] - SQ(dtau) * expdt;
  }
}

/* Fritsch & Butland (1984) centred derivative, rh/bezier_aux.c:34-51 */
static double cent_deriv(double dsup, double dsdn, double chiup, double chic,
                         double chidn)
{
  double fim1 = (chic - chiup) / dsup, fi = (chidn - chic) / dsdn, alpha;
  if (fim1 * fi > 0) {
    alpha = 0.333333333333333333333333 * (1.0 + dsdn / (dsdn + dsup));
    return (fim1 * fi) / ((1.0 - alpha) * fim1 + alpha * fi);
  }
  return 0.0;
}

/* Integration weights of the cubic Bezier interpolant (de la Cruz
   Rodriguez & Piskunov 2013), rh/bezier_aux.c:333-359 */
void rho_bezier3_coeffs(double dt, double *o)
{
  double dt2 = dt * dt, dt3 = dt2 * dt, dt4, eps;
  if (dt >= 5.e-2) {
    eps = exp(-dt);
    o[0] = (-6.0 + 6.0 * dt - 3.0 * dt2 + dt3 + 6.0 * eps) / dt3;
    dt3 = 1.0 / dt3;
    o[1] = (6.0 + (-6.0 - dt * (6.0 + dt * (3.0 + dt))) * eps) * dt3;
    o[2] = 3.0 * (6.0 + (-4.0 + dt) * dt - 2.0 * (3.0 + dt) * eps) * dt3;
    o[3] = 3.0 * (eps * (6.0 + dt2 + 4.0 * dt) + 2.0 * dt - 6.0) * dt3;
  } else {
    dt4 = dt2 * dt2;
    eps = 1.0 - dt + 0.5 * dt2 - dt3 / 6.0 + dt4 / 24.0;
    o[0] = 0.25 * dt - 0.05 * dt2 + dt3 / 120.0 - dt4 / 840.0;
    o[1] = 0.25 * dt - 0.20 * dt2 + dt3 / 12.0 - dt4 / 42.0;
    o[2] = 0.25 * dt - 0.10 * dt2 + dt3 * 0.025 - dt4 / 210.0;
    o[3] = 0.25 * dt - 0.15 * dt2 + dt3 * 0.05 - dt4 / 84.0;
  }
  o[4] = eps;
}

/* 4x4 inverse by cofactors (Cramer's rule), with the cofactor pairing of
   rh/bezier_aux.c:235-328 (Intel AP-928 "Streaming SIMD Extensions -
   Inverse of 4x4 Matrix", scalar form).  mat is row-major [4][4]. */
#define CRAMER_BODY(T)                                                    \
  T tmp[12], src[16], dst[16], det;                                       \
  int i;                                                                  \
  for (i = 0; i < 4; i++) {                                               \
    src[i] = mat[i * 4]; src[i + 4] = mat[i * 4 + 1];                     \
    src[i + 8] = mat[i * 4 + 2]; src[i + 12] = mat[i * 4 + 3];            \
  }                                                                       \
  tmp[0] = src[10] * src[15]; tmp[1] = src[11] * src[14];                 \
  tmp[2] = src[9] * src[15];  tmp[3] = src[11] * src[13];                 \
  tmp[4] = src[9] * src[14];  tmp[5] = src[10] * src[13];                 \
  tmp[6] = src[8] * src[15];  tmp[7] = src[11] * src[12];                 \
  tmp[8] = src[8] * src[14];  tmp[9] = src[10] * src[12];                 \
  tmp[10] = src[8] * src[13]; tmp[11] = src[9] * src[12];                 \
  dst[0] = tmp[0]*src[5] + tmp[3]*src[6] + tmp[4]*src[7];                 \
  dst[0] -= tmp[1]*src[5] + tmp[2]*src[6] + tmp[5]*src[7];                \
  dst[1] = tmp[1]*src[4] + tmp[6]*src[6] + tmp[9]*src[7];                 \
  dst[1] -= tmp[0]*src[4] + tmp[7]*src[6] + tmp[8]*src[7];                \
  dst[2] = tmp[2]*src[4] + tmp[7]*src[5] + tmp[10]*src[7];                \
  dst[2] -= tmp[3]*src[4] + tmp[6]*src[5] + tmp[11]*src[7];               \
  dst[3] = tmp[5]*src[4] + tmp[8]*src[5] + tmp[11]*src[6];                \
  dst[3] -= tmp[4]*src[4] + tmp[9]*src[5] + tmp[10]*src[6];               \
  dst[4] = tmp[1]*src[1] + tmp[2]*src[2] + tmp[5]*src[3];                 \
  dst[4] -= tmp[0]*src[1] + tmp[3]*src[2] + tmp[4]*src[3];                \
  dst[5] = tmp[0]*src[0] + tmp[7]*src[2] + tmp[8]*src[3];                 \
  dst[5] -= tmp[1]*src[0] + tmp[6]*src[2] + tmp[9]*src[3];                \
  dst[6] = tmp[3]*src[0] + tmp[6]*src[1] + tmp[11]*src[3];                \
  dst[6] -= tmp[2]*src[0] + tmp[7]*src[1] + tmp[10]*src[3];               \
  dst[7] = tmp[4]*src[0] + tmp[9]*src[1] + tmp[10]*src[2];                \
  dst[7] -= tmp[5]*src[0] + tmp[8]*src[1] + tmp[11]*src[2];               \
  tmp[0] = src[2]*src[7];  tmp[1] = src[3]*src[6];                        \
  tmp[2] = src[1]*src[7];  tmp[3] = src[3]*src[5];                        \
  tmp[4] = src[1]*src[6];  tmp[5] = src[2]*src[5];                        \
  tmp[6] = src[0]*src[7];  tmp[7] = src[3]*src[4];                        \
  tmp[8] = src[0]*src[6];  tmp[9] = src[2]*src[4];                        \
  tmp[10] = src[0]*src[5]; tmp[11] = src[1]*src[4];                       \
  dst[8] = tmp[0]*src[13] + tmp[3]*src[14] + tmp[4]*src[15];              \
  dst[8] -= tmp[1]*src[13] + tmp[2]*src[14] + tmp[5]*src[15];             \
  dst[9] = tmp[1]*src[12] + tmp[6]*src[14] + tmp[9]*src[15];              \
  dst[9] -= tmp[0]*src[12] + tmp[7]*src[14] + tmp[8]*src[15];             \
  dst[10] = tmp[2]*src[12] + tmp[7]*src[13] + tmp[10]*src[15];            \
  dst[10] -= tmp[3]*src[12] + tmp[6]*src[13] + tmp[11]*src[15];           \
  dst[11] = tmp[5]*src[12] + tmp[8]*src[13] + tmp[11]*src[14];            \
  dst[11] -= tmp[4]*src[12] + tmp[9]*src[13] + tmp[10]*src[14];           \
  dst[12] = tmp[2]*src[10] + tmp[5]*src[11] + tmp[1]*src[9];              \
  dst[12] -= tmp[4]*src[11] + tmp[0]*src[9] + tmp[3]*src[10];             \
  dst[13] = tmp[8]*src[11] + tmp[0]*src[8] + tmp[7]*src[10];              \
  dst[13] -= tmp[6]*src[10] + tmp[9]*src[11] + tmp[1]*src[8];             \
  dst[14] = tmp[6]*src[9] + tmp[11]*src[11] + tmp[3]*src[8];              \
  dst[14] -= tmp[10]*src[11] + tmp[2]*src[8] + tmp[7]*src[9];             \
  dst[15] = tmp[10]*src[10] + tmp[4]*src[8] + tmp[9]*src[9];              \
  dst[15] -= tmp[8]*src[9] + tmp[11]*src[10] + tmp[5]*src[8];             \
  det = src[0]*dst[0] + src[1]*dst[1] + src[2]*dst[2] + src[3]*dst[3];    \
  det = 1.0 / det;                                                        \
  for (i = 0; i < 16; i++) mat[i] = dst[i] * det;

static void matinv_d(double *mat) { CRAMER_BODY(double) }
static void matinv_f(float *mat) { CRAMER_BODY(float) }

/* Md^-1 V0 with Md held in double or float (rh/bezier_aux.c:109-117) */
static void solve4(double Md[4][4], const double *V0, double *V1, int fp32)
{
  int i, k;
  if (fp32) {
    float f[16];
    for (i = 0; i < 16; i++) f[i] = (float) Md[i / 4][i % 4];
    matinv_f(f);
    for (i = 0; i < 4; i++) {
      V1[i] = 0.0;
      for (k = 0; k < 4; k++) V1[i] += ((double) f[i * 4 + k]) * V0[k];
    }
  } else {
    double d[16];
    for (i = 0; i < 16; i++) d[i] = Md[i / 4][i % 4];
    matinv_d(d);
    for (i = 0; i < 4; i++) {
      V1[i] = 0.0;
      for (k = 0; k < 4; k++) V1[i] += d[i * 4 + k] * V0[k];
    }
  }
}

/* reduced propagation matrix K' (rh/stokesopac.c:28-88); q[3] are
   chi_Q,U,V at this depth, rho[3] the magneto-optical chip_Q,U,V (NULL: the
   terms are off) */
static void stokesK(const double *q, const double *rho, double chi_I, double K[4][4])
{
  int i, j;
  for (j = 0; j < 4; j++) for (i = 0; i < 4; i++) K[j][i] = 0.0;
  K[0][1] = q[0]; K[0][2] = q[1]; K[0][3] = q[2];
  if (rho) { K[1][2] = rho[2]; K[1][3] = rho[1]; K[2][3] = rho[0]; }
  for (j = 0; j < 3; j++)
    for (i = j + 1; i < 4; i++) {
      K[j][i] /= chi_I;
      K[i][j] = K[j][i];
    }
  if (rho) { K[1][3] *= -1.0; K[2][1] *= -1.0; K[3][2] *= -1.0; }
}

/* m4m: c = b . a as rh/bezier_aux.c:94-103 accumulates it */
static void m4m(double a[4][4], double b[4][4], double c[4][4])
{
  int i, j, k;
  for (j = 0; j < 4; j++) for (i = 0; i < 4; i++) c[j][i] = 0.0;
  for (j = 0; j < 4; j++)
    for (i = 0; i < 4; i++)
      for (k = 0; k < 4; k++) c[j][i] += a[k][i] * b[j][k];
}

/* Cubic DELO-Bezier polarized solver for the ray leaving the top
   (to_obs), rh/rhf1d/bezier_1D.c:52-300.  S4/chiQUV/I4 are [4|3][n]. */
void rho_formal_stokes(int Ndep, const double *z, double mu, const double *chi,
                       const double *S4, const double *chiQUV, const double *chip,
                       const double *B2, int fp32, double *I4)
{
  static const double ident[4][4] = {{1,0,0,0},{0,1,0,0},{0,0,1,0},{0,0,0,1}};
  const int dk = -1, k_start = Ndep - 1, k_end = 0;
  double imu = 1.0 / mu, dtau_uw, dtau_dw = 0.0, c1, c2, w[3], dsdn2, dchi_dn,
         I_upw[4], dchi_up, dchi_c, dt03, dsup, dsdn, dt, co[5];
  double Ku[4][4], K0[4][4], Kd[4][4], dKu[4][4], dK0[4][4];
  double Su[4], S0[4], Sd[4], dSu[4], dS0[4];
  double A[4][4], Ma[4][4], Mb[4][4], Mc[4][4], Md[4][4], V0[4], V1[4], q[3], r[3];
  int k, n, m, i, j;
#define IQ(s, kk) I4[(s) * Ndep + (kk)]
#define SV(s, kk) S4[(s) * Ndep + (kk)]
#define LOADK(kk, KK) do { q[0] = chiQUV[kk]; q[1] = chiQUV[Ndep + (kk)]; \
    q[2] = chiQUV[2 * Ndep + (kk)]; \
    if (chip) { r[0] = chip[kk]; r[1] = chip[Ndep + (kk)]; r[2] = chip[2 * Ndep + (kk)]; } \
    stokesK(q, chip ? r : NULL, chi[kk], KK); } while (0)

  dtau_uw = 0.5 * imu * (chi[k_start] + chi[k_start + dk]) *
            fabs(z[k_start] - z[k_start + dk]);
  if (B2) I_upw[0] = B2[1] - (B2[0] - B2[1]) / dtau_uw;
  else I_upw[0] = 0.0;
  for (n = 1; n < 4; n++) I_upw[n] = 0.0;
  for (n = 0; n < 4; n++) IQ(n, k_start) = I_upw[n];

  k = k_start + dk;
  dsup = fabs(z[k] - z[k - dk]) * imu;
  dsdn = fabs(z[k + dk] - z[k]) * imu;
  dchi_up = (chi[k] - chi[k - dk]) / dsup;
  dchi_c = cent_deriv(dsup, dsdn, chi[k - dk], chi[k], chi[k + dk]);
  c2 = MAXV(chi[k] - (dsup / 3.0) * dchi_c, 0.0);
  c1 = MAXV(chi[k - dk] + (dsup / 3.0) * dchi_up, 0.0);
  dtau_uw = 0.25 * dsup * (chi[k] + chi[k - dk] + c1 + c2);

  LOADK(k_start, Ku);
  LOADK(k_start + dk, K0);
  for (n = 0; n < 4; n++) { Su[n] = SV(n, k_start); S0[n] = SV(n, k_start + dk); }
  for (n = 0; n < 4; n++) {
    dSu[n] = (S0[n] - Su[n]) / dtau_uw;
    for (m = 0; m < 4; m++) dKu[n][m] = (K0[n][m] - Ku[n][m]) / dtau_uw;
  }

  for (k = k_start + dk; k != k_end; k += dk) {
    dsdn = fabs(z[k + dk] - z[k]) * imu;
    if (fabs(k - k_end) > 1) {
      dsdn2 = fabs(z[k + 2 * dk] - z[k + dk]) * imu;
      dchi_dn = cent_deriv(dsdn, dsdn2, chi[k], chi[k + dk], chi[k + 2 * dk]);
    } else
      dchi_dn = (chi[k + dk] - chi[k]) / dsdn;
    c2 = MAXV(chi[k] + (dsdn / 3.0) * dchi_c, 0.0);
    c1 = MAXV(chi[k + dk] - (dsdn / 3.0) * dchi_dn, 0.0);
    dtau_dw = 0.25 * dsdn * (chi[k] + chi[k + dk] + c1 + c2);
    dt = dtau_uw; dt03 = dt / 3.0;
    rho_bezier3_coeffs(dt, co);

    LOADK(k + dk, Kd);
    for (n = 0; n < 4; n++) Sd[n] = SV(n, k + dk);
    for (j = 0; j < 4; j++)
      for (i = 0; i < 4; i++)
        dK0[j][i] = cent_deriv(dtau_uw, dtau_dw, Ku[j][i], K0[j][i], Kd[j][i]);
    for (i = 0; i < 4; i++) dS0[i] = cent_deriv(dtau_uw, dtau_dw, Su[i], S0[i], Sd[i]);

    m4m(Ku, Ku, Ma);
    m4m(K0, K0, A);
    for (j = 0; j < 4; j++)
      for (i = 0; i < 4; i++) {
        Md[j][i] = ident[j][i] + co[0] * K0[j][i] - co[2] *
          -(dt03 * (A[j][i] + dK0[j][i] + K0[j][i]) + K0[j][i]);
        Ma[j][i] = co[4] * ident[j][i] - co[1] * Ku[j][i] + co[3] *
          (dt03 * (Ma[j][i] + dKu[j][i] + Ku[j][i]) - Ku[j][i]);
        Mb[j][i] = co[1] * ident[j][i] + co[3] * (ident[j][i] - dt03 * Ku[j][i]);
        Mc[j][i] = co[0] * ident[j][i] + co[2] * (ident[j][i] + dt03 * K0[j][i]);
      }
    for (i = 0; i < 4; i++) {
      V0[i] = 0.0;
      for (j = 0; j < 4; j++)
        V0[i] += Ma[i][j] * IQ(j, k - dk) + Mb[i][j] * Su[j] + Mc[i][j] * S0[j];
      V0[i] += dt03 * (co[2] * dS0[i] - co[3] * dSu[i]);
    }
    solve4(Md, V0, V1, fp32);
    for (i = 0; i < 4; i++) IQ(i, k) = V1[i];

    memcpy(Su, S0, sizeof(Su));
    memcpy(S0, Sd, sizeof(S0));
    memcpy(dSu, dS0, sizeof(dSu));
    memcpy(Ku, K0, sizeof(Ku));
    memcpy(K0, Kd, sizeof(K0));
    memcpy(dKu, dK0, sizeof(dKu));
    dtau_uw = dtau_dw;
    dsup = dsdn;
    dchi_up = dchi_c;
    dchi_c = dchi_dn;
  }

  k = k_end;
  dtau_uw = 0.5 * imu * (chi[k] + chi[k - dk]) * fabs(z[k] - z[k - dk]);
  rho_w3(dtau_uw, w);
  for (n = 0; n < 4; n++) V0[n] = w[0] * SV(n, k) + w[1] * -dSu[n];
  for (n = 0; n < 4; n++) {
    for (m = 0; m < 4; m++) {
      A[n][m] = -w[1] / dtau_uw * Ku[n][m];
      Md[n][m] = (w[0] - w[1] / dtau_uw) * K0[n][m];
    }
    A[n][n] = 1.0 - w[0];
    Md[n][n] = 1.0;
  }
  for (n = 0; n < 4; n++)
    for (m = 0; m < 4; m++) V0[n] += A[n][m] * IQ(m, k - dk);
  solve4(Md, V0, V1, fp32);
  for (n = 0; n < 4; n++) IQ(n, k) = V1[n];
#undef IQ
#undef SV
#undef LOADK
}

/* Cubic Bezier solver for unpolarized light, ray leaving the top,
   rh/rhf1d/bezier_1D.c:306-541 (without the log gf derivatives) */
void rho_formal_bezier(int Ndep, const double *z, double mu, const double *chi,
                       const double *S, const double *B2, double *I)
{
  const int dk = -1, k_start = Ndep - 1, k_end = 0;
  double zmu = 1.0 / mu, dtau_uw, dtau_dw = 0.0, dS_uw, I_upw, c1, c2, w[3];
  double dsup, dsdn, dt03, co[5], dS_up, dS_c, dchi_up, dchi_c, dchi_dn = 0.0, dsdn2;
  int k;

  dtau_uw = 0.5 * zmu * (chi[k_start] + chi[k_start + dk]) *
            fabs(z[k_start] - z[k_start + dk]);
  I_upw = B2 ? B2[1] - (B2[0] - B2[1]) / dtau_uw : 0.0;
  I[k_start] = I_upw;

  k = k_start + dk;
  dsup = fabs(z[k] - z[k - dk]) * zmu;
  dsdn = fabs(z[k + dk] - z[k]) * zmu;
  dchi_up = (chi[k] - chi[k - dk]) / dsup;
  dchi_c = cent_deriv(dsup, dsdn, chi[k - dk], chi[k], chi[k + dk]);
  c1 = MAXV(chi[k] - (dsup / 3.0) * dchi_c, 0.0);
  c2 = MAXV(chi[k - dk] + (dsup / 3.0) * dchi_up, 0.0);
  dtau_uw = dsup * (chi[k] + chi[k - dk] + c1 + c2) * 0.25;
  dS_up = (S[k] - S[k - dk]) / dtau_uw;

  for (k = k_start + dk; k != k_end + dk; k += dk) {
    if (k != k_end) {
      dsdn = fabs(z[k + dk] - z[k]) * zmu;
      if (fabs(k - k_end) > 1) {
        dsdn2 = fabs(z[k + 2 * dk] - z[k + dk]) * zmu;
        dchi_dn = cent_deriv(dsdn, dsdn2, chi[k], chi[k + dk], chi[k + 2 * dk]);
      } else
        dchi_dn = (chi[k + dk] - chi[k]) / dsdn;
      c1 = MAXV(chi[k] + (dsdn / 3.0) * dchi_c, 0.0);
      c2 = MAXV(chi[k + dk] - (dsdn / 3.0) * dchi_dn, 0.0);
      dtau_dw = dsdn * (chi[k] + chi[k + dk] + c1 + c2) * 0.25;
      dt03 = dtau_uw / 3.0;
      rho_bezier3_coeffs(dtau_uw, co);
      dS_c = cent_deriv(dtau_uw, dtau_dw, S[k - dk], S[k], S[k + dk]);
      c1 = MAXV(S[k] - dt03 * dS_c, 0.0);
      c2 = MAXV(S[k - dk] + dt03 * dS_up, 0.0);
      I[k] = I_upw * co[4] + co[0] * S[k] + co[1] * S[k - dk] + co[2] * c1 + co[3] * c2;
    } else {
      dtau_uw = 0.5 * zmu * (chi[k] + chi[k - dk]) * fabs(z[k] - z[k - dk]);
      dS_uw = -(S[k] - S[k - dk]) / dtau_uw;
      rho_w3(dtau_uw, w);
      I[k] = (1.0 - w[0]) * I_upw + w[0] * S[k] + w[1] * dS_uw;
    }
    I_upw = I[k];
    dsup = dsdn;
    dchi_up = dchi_c;
    dchi_c = dchi_dn;
    dtau_uw = dtau_dw;
    dS_up = dS_c;
  }
  (void) dsup;
  (void) dchi_up;
}

/* Scalar cubic Bezier along the inward ray (to_obs = 0: depth 0 to n-1)
   with zero incident intensity at the top (ZERO boundary,
   rh/rhf1d/multiatmos.c:62-66).  rh/rhf1d/bezier_1D.c:306-541 */
void rho_formal_bezier_in(int Ndep, const double *z, double mu, const double *chi,
                          const double *S, double Itop, double *I)
{
  const int dk = 1, k_start = 0, k_end = Ndep - 1;
  double zmu = 1.0 / mu, dtau_uw, dtau_dw = 0.0, dS_uw, I_upw, c1, c2, w[3];
  double dsup, dsdn, dt03, co[5], dS_up, dS_c, dchi_up, dchi_c, dchi_dn = 0.0, dsdn2;
  int k;

  /* ZERO or IRRADIATED top boundary (rh/rhf1d/bezier_1D.c:370-380) */
  I_upw = Itop;
  I[k_start] = I_upw;
  k = k_start + dk;
  dsup = fabs(z[k] - z[k - dk]) * zmu;
  dsdn = fabs(z[k + dk] - z[k]) * zmu;
  dchi_up = (chi[k] - chi[k - dk]) / dsup;
  dchi_c = cent_deriv(dsup, dsdn, chi[k - dk], chi[k], chi[k + dk]);
  c1 = MAXV(chi[k] - (dsup / 3.0) * dchi_c, 0.0);
  c2 = MAXV(chi[k - dk] + (dsup / 3.0) * dchi_up, 0.0);
  dtau_uw = dsup * (chi[k] + chi[k - dk] + c1 + c2) * 0.25;
  dS_up = (S[k] - S[k - dk]) / dtau_uw;

  for (k = k_start + dk; k != k_end + dk; k += dk) {
    if (k != k_end) {
      dsdn = fabs(z[k + dk] - z[k]) * zmu;
      if (abs(k - k_end) > 1) {
        dsdn2 = fabs(z[k + 2 * dk] - z[k + dk]) * zmu;
        dchi_dn = cent_deriv(dsdn, dsdn2, chi[k], chi[k + dk], chi[k + 2 * dk]);
      } else
        dchi_dn = (chi[k + dk] - chi[k]) / dsdn;
      c1 = MAXV(chi[k] + (dsdn / 3.0) * dchi_c, 0.0);
      c2 = MAXV(chi[k + dk] - (dsdn / 3.0) * dchi_dn, 0.0);
      dtau_dw = dsdn * (chi[k] + chi[k + dk] + c1 + c2) * 0.25;
      dt03 = dtau_uw / 3.0;
      rho_bezier3_coeffs(dtau_uw, co);
      dS_c = cent_deriv(dtau_uw, dtau_dw, S[k - dk], S[k], S[k + dk]);
      c1 = MAXV(S[k] - dt03 * dS_c, 0.0);
      c2 = MAXV(S[k - dk] + dt03 * dS_up, 0.0);
      I[k] = I_upw * co[4] + co[0] * S[k] + co[1] * S[k - dk] + co[2] * c1 + co[3] * c2;
    } else {
      dtau_uw = 0.5 * zmu * (chi[k] + chi[k - dk]) * fabs(z[k] - z[k - dk]);
      dS_uw = -(S[k] - S[k - dk]) / dtau_uw;
      rho_w3(dtau_uw, w);
      I[k] = (1.0 - w[0]) * I_upw + w[0] * S[k] + w[1] * dS_uw;
    }
    I_upw = I[k];
    dsup = dsdn;
    dchi_up = dchi_c;
    dchi_c = dchi_dn;
    dtau_uw = dtau_dw;
    dS_up = dS_c;
  }
  (void) dsup;
  (void) dchi_up;
}

/* Scalar cubic Bezier along the emergent ray with the reference's analytic
   d I / d log gf recursion (rh/rhf1d/bezier_1D.c:416-536):
   Z = (-dchi I + deta) / chi through the intensity's Bezier weights, where
   I is the one array the inward ray filled first (rh/rhf1d/formal.c:168-172),
   so the downwind I in Z is still the inward value.  On entry I holds the
   inward intensities; on return the emergent ones.  dchi, deta are
   [npar][n]; dI0 receives d I(top) / d p for each parameter. */
void rho_formal_bezier_rf(int Ndep, const double *z, double mu, const double *chi,
                          const double *S, const double *B2, int npar,
                          const double *dchi, const double *deta, double *I, double *dI0)
{
  const int dk = -1, k_start = Ndep - 1, k_end = 0;
  double zmu = 1.0 / mu, dtau_uw, dtau_dw = 0.0, dS_uw, I_upw, c1, c2, w[3];
  double dsup, dsdn, dt03, co[5], dS_up, dS_c, dchi_up, dchi_c, dchi_dn = 0.0, dsdn2;
  double Zk, Zkm1, Zkp1, *dZk, *dZup, *dI_upw, *dI;
  int k, p;
#define DCHI(p, k) dchi[(size_t) (p) * Ndep + (k)]
#define DETA(p, k) deta[(size_t) (p) * Ndep + (k)]

  dZk = calloc((size_t) 4 * npar + 1, sizeof(double));
  dZup = dZk + npar;
  dI_upw = dZup + npar;
  dI = dI_upw + npar;
  dtau_uw = 0.5 * zmu * (chi[k_start] + chi[k_start + dk]) *
            fabs(z[k_start] - z[k_start + dk]);
  I_upw = B2[1] - (B2[0] - B2[1]) / dtau_uw;
  I[k_start] = I_upw;

  k = k_start + dk;
  dsup = fabs(z[k] - z[k - dk]) * zmu;
  dsdn = fabs(z[k + dk] - z[k]) * zmu;
  dchi_up = (chi[k] - chi[k - dk]) / dsup;
  dchi_c = cent_deriv(dsup, dsdn, chi[k - dk], chi[k], chi[k + dk]);
  c1 = MAXV(chi[k] - (dsup / 3.0) * dchi_c, 0.0);
  c2 = MAXV(chi[k - dk] + (dsup / 3.0) * dchi_up, 0.0);
  dtau_uw = dsup * (chi[k] + chi[k - dk] + c1 + c2) * 0.25;
  dS_up = (S[k] - S[k - dk]) / dtau_uw;
  for (p = 0; p < npar; p++) {
    dI_upw[p] = 0.0;
    Zk = -DCHI(p, k) / chi[k] * I[k] + DETA(p, k) / chi[k];
    Zkm1 = -DCHI(p, k - dk) / chi[k - dk] * I[k - dk] + DETA(p, k - dk) / chi[k - dk];
    dZup[p] = (Zk - Zkm1) / dtau_uw;
  }

  for (k = k_start + dk; k != k_end + dk; k += dk) {
    if (k != k_end) {
      dsdn = fabs(z[k + dk] - z[k]) * zmu;
      if (abs(k - k_end) > 1) {
        dsdn2 = fabs(z[k + 2 * dk] - z[k + dk]) * zmu;
        dchi_dn = cent_deriv(dsdn, dsdn2, chi[k], chi[k + dk], chi[k + 2 * dk]);
      } else
        dchi_dn = (chi[k + dk] - chi[k]) / dsdn;
      c1 = MAXV(chi[k] + (dsdn / 3.0) * dchi_c, 0.0);
      c2 = MAXV(chi[k + dk] - (dsdn / 3.0) * dchi_dn, 0.0);
      dtau_dw = dsdn * (chi[k] + chi[k + dk] + c1 + c2) * 0.25;
      dt03 = dtau_uw / 3.0;
      rho_bezier3_coeffs(dtau_uw, co);
      dS_c = cent_deriv(dtau_uw, dtau_dw, S[k - dk], S[k], S[k + dk]);
      c1 = MAXV(S[k] - dt03 * dS_c, 0.0);
      c2 = MAXV(S[k - dk] + dt03 * dS_up, 0.0);
      I[k] = I_upw * co[4] + co[0] * S[k] + co[1] * S[k - dk] + co[2] * c1 + co[3] * c2;
      for (p = 0; p < npar; p++) {
        Zk = -DCHI(p, k) * I[k] + DETA(p, k);
        Zk /= chi[k];
        Zkm1 = -DCHI(p, k - dk) * I[k - dk] + DETA(p, k - dk);
        Zkm1 /= chi[k - dk];
        Zkp1 = -DCHI(p, k + dk) * I[k + dk] + DETA(p, k + dk);
        Zkp1 /= chi[k + dk];
        dZk[p] = cent_deriv(dtau_uw, dtau_dw, Zkm1, Zk, Zkp1);
        c1 = MAXV(Zk - dt03 * dZk[p], 0.0);
        c2 = MAXV(Zkm1 + dt03 * dZup[p], 0.0);
        dI[p] = dI_upw[p] * co[4] + co[0] * Zk + co[1] * Zkm1 + co[2] * c1 + co[3] * c2;
      }
    } else {
      dtau_uw = 0.5 * zmu * (chi[k] + chi[k - dk]) * fabs(z[k] - z[k - dk]);
      dS_uw = -(S[k] - S[k - dk]) / dtau_uw;
      rho_w3(dtau_uw, w);
      I[k] = (1.0 - w[0]) * I_upw + w[0] * S[k] + w[1] * dS_uw;
      for (p = 0; p < npar; p++) {
        Zk = DCHI(p, k) / chi[k] * I[k] - DETA(p, k) / chi[k];
        Zkm1 = DCHI(p, k - dk) / chi[k - dk] * I[k - dk] - DETA(p, k - dk) / chi[k - dk];
        dZk[p] = -(Zk - Zkm1) / dtau_uw;
        dI[p] = (1.0 - w[0]) * dI_upw[p] + w[0] * Zk + w[1] * dZk[p];
      }
    }
    I_upw = I[k];
    dsup = dsdn;
    dchi_up = dchi_c;
    dchi_c = dchi_dn;
    dtau_uw = dtau_dw;
    dS_up = dS_c;
    for (p = 0; p < npar; p++) {
      dI_upw[p] = dI[p];
      dZup[p] = dZk[p];
    }
  }
  for (p = 0; p < npar; p++) dI0[p] = dI_upw[p];
  free(dZk);
  (void) dsup;
  (void) dchi_up;
#undef DCHI
#undef DETA
}

/* Feautrier solution, STANDARD order, ZERO upper boundary, thermalised
   lower boundary; returns the emergent intensity.  rh/rhf1d/feautrier.c:56-202 */
static double formal_feautrier_h0(int Ndep, const double *z, double mu, const double *chi,
                                  const double *S, const double *B2, double h0);

double rho_formal_feautrier(int Ndep, const double *z, double mu,
                            const double *chi, const double *S, const double *B2)
{
  return formal_feautrier_h0(Ndep, z, mu, chi, S, B2, 0.0);
}

/* h0: the intensity irradiating the top (rh/rhf1d/feautrier.c:84-97) */
static double formal_feautrier_h0(int Ndep, const double *z, double mu, const double *chi,
                                  const double *S, const double *B2, double h0)
{
  double r0 = 0.0, rN = 0.0, hN, f0, fN, dtau_mid, zmu = 0.5 / mu;
  double *dtau = malloc(8 * (size_t) Ndep), *abc = malloc(8 * (size_t) Ndep),
         *A1 = malloc(8 * (size_t) Ndep), *C1 = malloc(8 * (size_t) Ndep),
         *F = malloc(8 * (size_t) Ndep), *ztmp = malloc(8 * (size_t) Ndep),
         *Stmp = malloc(8 * (size_t) Ndep), *P = malloc(8 * (size_t) Ndep), Iplus;
  int k;
  for (k = 0; k < Ndep - 1; k++)
    dtau[k] = zmu * (chi[k] + chi[k + 1]) * (z[k] - z[k + 1]);
  f0 = (1.0 - r0) / (1.0 + r0);
  abc[0] = 1.0 + 2.0 * f0 / dtau[0];
  C1[0] = 2.0 / SQ(dtau[0]);
  Stmp[0] = S[0] + 2.0 * h0 / ((1.0 + r0) * dtau[0]);
  hN = B2 ? B2[1] - (B2[0] - B2[1]) / dtau[Ndep - 2] : 0.0;
  fN = (1.0 - rN) / (1.0 + rN);
  abc[Ndep - 1] = 1.0 + 2.0 * fN / dtau[Ndep - 2];
  A1[Ndep - 1] = 2.0 / SQ(dtau[Ndep - 2]);
  Stmp[Ndep - 1] = S[Ndep - 1] + 2.0 * hN / ((1.0 + rN) * dtau[Ndep - 2]);
  for (k = 1; k < Ndep - 1; k++) {
    dtau_mid = 0.5 * (dtau[k] + dtau[k - 1]);
    A1[k] = 1.0 / (dtau_mid * dtau[k - 1]);
    C1[k] = 1.0 / (dtau_mid * dtau[k]);
    abc[k] = 1.0;
    Stmp[k] = S[k];
  }
  F[0] = abc[0] / C1[0];
  ztmp[0] = Stmp[0] / (abc[0] + C1[0]);
  for (k = 1; k < Ndep - 1; k++) {
    F[k] = (abc[k] + A1[k] * F[k - 1] / (1.0 + F[k - 1])) / C1[k];
    ztmp[k] = (Stmp[k] + A1[k] * ztmp[k - 1]) / (C1[k] * (1.0 + F[k]));
  }
  P[Ndep - 1] = (Stmp[Ndep - 1] + A1[Ndep - 1] * ztmp[Ndep - 2]) /
    (abc[Ndep - 1] + A1[Ndep - 1] * (F[Ndep - 2] / (1.0 + F[Ndep - 2])));
  for (k = Ndep - 2; k >= 0; k--) P[k] = P[k + 1] / (1.0 + F[k]) + ztmp[k];
  Iplus = (1.0 + f0) * P[0] - h0 / (1.0 + r0);
  free(dtau); free(abc); free(A1); free(C1); free(F); free(ztmp); free(Stmp); free(P);
  return Iplus;
}

/* ------------------------------------------------------------------ */
/* equilibria                                                          */

/* Saha-Boltzmann populations of one model atom at one depth,
   rh/ltepops.c:33-113 (Debye lowering disabled there) */
static void ltepops(const rhg_model_t *m, int ia, double T, double ne,
                    double ntotal, double *nstar)
{
  int i0 = m->atom_level_ptr[ia], nl = m->atom_level_ptr[ia + 1] - i0, i, mm, dZ;
  const double *E = m->level_E + i0, *g = m->level_g + i0;
  const int32_t *st = m->level_stage + i0;
  double c1 = (HPLANCK / (2.0 * PI * M_ELECTRON)) * (HPLANCK / KBOLTZMANN);
  double cNe_T = 0.5 * ne * pow(c1 / T, 1.5), sum = 1.0, dE, gi0, dE_kT;
  for (i = 1; i < nl; i++) {
    dE = E[i] - E[0];
    gi0 = g[i] / g[0];
    dZ = st[i] - st[0];
    dE_kT = dE / (KBOLTZMANN * T);
    nstar[i] = gi0 * exp(-dE_kT);
    for (mm = 1; mm <= dZ; mm++) nstar[i] /= cNe_T;
    sum += nstar[i];
  }
  nstar[0] = ntotal / sum;
  for (i = 1; i < nl; i++) nstar[i] *= nstar[0];
}

/* equilibrium constant of a molecule, rh/chemequil.c:452-533 */
static double equilconstant(const rhg_model_t *m, int im, double T)
{
  int i, mk, neqc = m->mol_neqc[im];
  const double *c = m->mol_eqc + 8 * im;
  double kT, t, eqc = 0.0, theta, cgs_to_SI = 1.0;
  if (T < m->mol_tmin[im] || T > m->mol_tmax[im]) return 0.0;
  switch (m->mol_fit[im]) {
  case 0: /* KURUCZ_70 */
  case 1: /* KURUCZ_85 */
    t = (m->mol_fit[im] == 1) ? T * 1.0E-4 : T;
    kT = KBOLTZMANN * T;
    mk = m->mol_nnuclei[im] - 1 - m->mol_charge[im];
    eqc = c[0];
    for (i = 1; i < neqc; i++) eqc = eqc * t + c[i];
    eqc = exp(m->mol_ediss[im] / kT + eqc - 1.5 * mk * log(T));
    cgs_to_SI = pow(CUBE(CM_TO_M), mk);
    break;
  case 3: /* IRWIN_81 uses the SAUVAL_TATUM_84 form */
  case 2:
    theta = THETA0 / T;
    t = log10(theta);
    kT = KBOLTZMANN * T;
    eqc = c[0];
    for (i = 1; i < neqc; i++) eqc = eqc * t + c[i];
    eqc = POW10((m->mol_ediss[im] / EV) * theta - eqc) * kT;
    cgs_to_SI = 1.0;
    break;
  case 4: /* TSUJI_73 */
    theta = THETA0 / T;
    kT = KBOLTZMANN * T;
    eqc = c[0];
    for (i = 1; i < neqc; i++) eqc = eqc * theta + c[i];
    eqc = SQ(kT) * POW10(-eqc);
    cgs_to_SI = pow(CUBE(CM_TO_M) / ERG_TO_JOULE, m->mol_nnuclei[im] - 1);
    break;
  }
  return eqc * cgs_to_SI;
}

/* Crout LU with implicit pivoting and one step of iterative improvement,
   rh/ludcmp.c:36-177 (Press et al., Numerical Recipes) */
#define NEQ_MAX 48
static void lu_decomp(int N, double A[][NEQ_MAX], int *index)
{
  int i, j, k, imax = 0;
  double big, dum, sum, temp, vv[NEQ_MAX];
  for (i = 0; i < N; i++) {
    big = 0.0;
    for (j = 0; j < N; j++) if ((temp = fabs(A[i][j])) > big) big = temp;
    vv[i] = 1.0 / big;
  }
  for (j = 0; j < N; j++) {
    for (i = 0; i < j; i++) {
      sum = A[i][j];
      for (k = 0; k < i; k++) sum -= A[i][k] * A[k][j];
      A[i][j] = sum;
    }
    big = 0.0;
    for (i = j; i < N; i++) {
      sum = A[i][j];
      for (k = 0; k < j; k++) sum -= A[i][k] * A[k][j];
      A[i][j] = sum;
      if ((dum = vv[i] * fabs(sum)) >= big) { big = dum; imax = i; }
    }
    if (j != imax) {
      for (k = 0; k < N; k++) { dum = A[imax][k]; A[imax][k] = A[j][k]; A[j][k] = dum; }
      vv[imax] = vv[j];
    }
    index[j] = imax;
    if (A[j][j] == 0.0) A[j][j] = 1.0e-20;
    if (j != N) {
      dum = 1.0 / A[j][j];
      for (i = j + 1; i < N; i++) A[i][j] *= dum;
    }
  }
}

static void lu_backsubst(int N, double A[][NEQ_MAX], const int *index, double *b)
{
  int i, j, ii = -1, ip;
  double sum;
  for (i = 0; i < N; i++) {
    ip = index[i];
    sum = b[ip];
    b[ip] = b[i];
    if (ii >= 0) { for (j = ii; j < i; j++) sum -= A[i][j] * b[j]; }
    else if (sum) ii = i;
    b[i] = sum;
  }
  for (i = N - 1; i >= 0; i--) {
    sum = b[i];
    for (j = i + 1; j < N; j++) sum -= A[i][j] * b[j];
    b[i] = sum / A[i][i];
  }
}

static void solve_linear_improve(int N, double A[][NEQ_MAX], double *b)
{
  double Ac[NEQ_MAX][NEQ_MAX], bc[NEQ_MAX], res[NEQ_MAX];
  int index[NEQ_MAX], i, j;
  for (i = 0; i < N; i++) { bc[i] = b[i]; for (j = 0; j < N; j++) Ac[i][j] = A[i][j]; }
  lu_decomp(N, A, index);
  lu_backsubst(N, A, index, b);
  for (i = 0; i < N; i++) {
    res[i] = bc[i];
    for (j = 0; j < N; j++) res[i] -= Ac[i][j] * b[j];
  }
  lu_backsubst(N, A, index, res);
  for (i = 0; i < N; i++) b[i] += res[i];
}

/* Chemical equilibrium at one depth, rh/chemequil.c:107-393 (Ng order 0,
   so Accelerate only stores the solution; MaxChange over the last two).
   nstar_all holds all level populations of all model atoms (offsets
   atom_level_ptr) and is rescaled in place for atoms bound in molecules;
   ntotal[atom] likewise.  Returns nHmin; nmol_out gets molecule numbers. */
/* iterations of chemeq on this thread (work census of rho_hse) */
static _Thread_local long long g_chem_iter = 0;

static double chemeq(const rhg_model_t *m, double T, double ne, double nHtot,
                     double *nstar_all, double *ntotal, double *nmol_out)
{
  const double CI = (HPLANCK / (2.0 * PI * M_ELECTRON)) * (HPLANCK / KBOLTZMANN);
  int Nn = m->nnuclei, Nm = m->nmol, N = Nn + Nm, i, j, nu, niter, ia;
  double f[NEQ_MAX], n[NEQ_MAX], a[NEQ_MAX], df[NEQ_MAX][NEQ_MAX], Phi[NEQ_MAX],
         fn0[NEQ_MAX], prev[2][NEQ_MAX], PhiHmin, fHmin, saha, dnmax = 0.0, fraction;
  int count;
  for (i = 0; i < N; i++) a[i] = 0.0;     /* calloc'd in the reference */
  for (i = 0; i < Nn; i++) {
    ia = m->nucl_atom[i];
    {
      int l0 = m->atom_level_ptr[ia], nl = m->atom_level_ptr[ia + 1] - l0;
      fn0[i] = 0.0;
      for (j = 0; j < nl; j++) {
        if (m->level_stage[l0 + j] > 0) break;
        fn0[i] += nstar_all[l0 + j];
      }
      fn0[i] /= ntotal[ia];
      a[i] = m->atom_abund[ia] * nHtot;
    }
  }
  PhiHmin = 0.25 * pow(CI / T, 1.5) * exp(0.754 * EV / (KBOLTZMANN * T));
  fHmin = ne * fn0[0] * PhiHmin;
  for (i = 0; i < Nm; i++) Phi[i] = equilconstant(m, i, T);
  for (i = 0; i < Nn; i++) n[i] = a[i];
  for (i = 0; i < Nm; i++) n[Nn + i] = 0.0;
  count = 1;
  for (i = 0; i < N; i++) prev[0][i] = n[i];

  niter = 1;
  while (niter <= N_MAX_CHEM_ITER) {
    for (i = 0; i < N; i++) {
      f[i] = n[i] - a[i];
      for (j = 0; j < N; j++) df[i][j] = 0.0;
      df[i][i] = 1.0;
    }
    f[0] += fHmin * n[0];
    df[0][0] += fHmin;
    for (i = 0; i < Nm; i++) {
      int nel = m->mol_nel[i];
      saha = Phi[i];
      for (j = 0; j < nel; j++) {
        nu = m->mol_el_nucl[4 * i + j];
        saha *= pow(fn0[nu] * n[nu], m->mol_el_count[4 * i + j]);
        f[nu] += m->mol_el_count[4 * i + j] * n[Nn + i];
      }
      saha /= pow(ne, m->mol_charge[i]);
      f[Nn + i] -= saha;
      for (j = 0; j < nel; j++) {
        nu = m->mol_el_nucl[4 * i + j];
        df[nu][Nn + i] += m->mol_el_count[4 * i + j];
        df[Nn + i][nu] = -saha * (m->mol_el_count[4 * i + j] / n[nu]);
      }
    }
    solve_linear_improve(N, df, f);
    for (i = 0; i < N; i++) n[i] -= f[i];
    /* Accelerate with order 0: store, count++ */
    for (i = 0; i < N; i++) prev[count % 2][i] = n[i];
    count++;
    {
      const double *old = prev[(count - 2) % 2], *nw = prev[(count - 1) % 2];
      dnmax = 0.0;
      for (i = 0; i < N; i++)
        if (nw[i]) dnmax = MAXV(dnmax, fabs((nw[i] - old[i]) / nw[i]));
    }
    g_chem_iter++;
    if (dnmax <= CHEM_ITER_LIMIT) break;
    niter++;
  }
  for (i = 0; i < Nn; i++) {
    ia = m->nucl_atom[i];
    {
      int l0 = m->atom_level_ptr[ia], nl = m->atom_level_ptr[ia + 1] - l0;
      fraction = n[i] / ntotal[ia];
      for (j = 0; j < nl; j++) nstar_all[l0 + j] *= fraction;
      ntotal[ia] = n[i];
    }
  }
  for (i = 0; i < Nm; i++) nmol_out[i] = n[Nn + i];
  return ne * (n[0] * PhiHmin);
}

/* Element stage populations from Kurucz partition functions,
   rh/ltepops.c:116-164 (one depth); nst[Nstage] */
static void ltepops_elem(const rhg_model_t *m, int ie, double T, double ne,
                         double nHtot, double *nst)
{
  const double C1 = (HPLANCK / (2.0 * PI * M_ELECTRON)) * (HPLANCK / KBOLTZMANN);
  int ns = m->el_nstage[ie], i, npf = m->npf;
  const double *lp = m->el_logpf + m->el_pf_ptr[ie];
  const double *ip = m->el_ionpot + m->el_pf_ptr[ie];
  double CT_ne = 2.0 * pow(C1 / T, -1.5) / ne, sum = 1.0, Uk, Ukp1;
  nst[0] = 1.0;
  Uk = linear1(npf, m->Tpf, lp, T);
  for (i = 1; i < ns; i++) {
    Ukp1 = linear1(npf, m->Tpf, lp + i * npf, T);
    nst[i] = nst[i - 1] * CT_ne * exp(Ukp1 - Uk - ip[i - 1] / (KBOLTZMANN * T));
    sum += nst[i];
    Uk = Ukp1;
  }
  nst[0] = m->el_abund[ie] * nHtot / sum;
  for (i = 1; i < ns; i++) nst[i] *= nst[0];
}

/* ------------------------------------------------------------------ */
/* one column                                                          */

typedef struct {
  int Ndep;
  double *T, *ne, *vz, *vturb, *B, *nHtot, *cosg, *cos2x, *sin2x;
  double *tau_ref, *cmass, *height;
  double *nstar;        /* [Ndep][nlevel] */
  double *ntotal;       /* [Ndep][natom]  */
  double *nHmin, *nH2, *nOH, *nCH, *nH0, *np, *npstar;
  double *nel;          /* [Ndep][nel][NSTAGE_MAX] */
  int moving;
} column_t;

#define NSTAGE_MAX 8

static void background_continuum(const rhg_model_t *m, const column_t *c, int ls,
                                 double *chi_ai, double *eta_ai)
{
  /* rh/background.c:329-465 (scattering terms omitted: in LTE they enter
     neither the opacity nor the source function, background.c:462-465) */
  const double *lc = m->lamconst + (size_t) ls * RHG_LC_NCONST;
  const double lambda = lc[RHG_LC_LAMBDA];
  const int nlev = m->atom_level_ptr[m->natom];
  double C0 = SQ(Q_ELECTRON) / (4.0 * PI * EPSILON_0) / sqrt(M_ELECTRON);
  double sigma_ff = 4.0 / 3.0 * sqrt(2.0 * PI / (3.0 * KBOLTZMANN)) * CUBE(C0) /
                    (HPLANCK * CLIGHT);
  int k, cp0 = m->lam_cont_ptr[ls], cp1 = m->lam_cont_ptr[ls + 1], e;
  int hasH = 0, hasM = 0;
  for (e = cp0; e < cp1; e++) {
    if (m->cont_atom[m->lam_cont_idx[e]] == 0) hasH = 1; else hasM = 1;
  }
  (void) hasM;
  for (k = 0; k < c->Ndep; k++) {
    double T = c->T[k], ne = c->ne[k], Bnu, chi, eta, stim, x;
    const double *ns = c->nstar + (size_t) k * nlev;
    chi_ai[k] = 0.0;
    eta_ai[k] = 0.0;
    {
      double hc_Tkla = lc[RHG_LC_HC_KLA] / T;
      Bnu = (hc_Tkla <= MAX_EXPONENT) ? lc[RHG_LC_TWOHNU3_C2] / (exp(hc_Tkla) - 1.0) : 0.0;
    }
    /* H- bound-free, rh/hydrogen.c:310-364 */
    if (lc[RHG_LC_HMIN_BF_ON] != 0.0) {
      stim = exp(-lc[RHG_LC_HC_KLA] / T);
      chi_ai[k] += c->nHmin[k] * (1.0 - stim) * lc[RHG_LC_HMIN_BF_ALPHA];
      eta_ai[k] += c->nHmin[k] * lc[RHG_LC_TWOHNU3_C2] * stim * lc[RHG_LC_HMIN_BF_ALPHA];
    }
    /* H- free-free, rh/hydrogen.c:370-548 */
    if (lc[RHG_LC_HMIN_FF_LIDX] >= 0.0) {
      double ti = clamp_index(NTHETA, m->hminff_theta, THETA0 / T);
      double pe = ne * KBOLTZMANN * T;
      double kappa = bilinear(NTHETA, NFF, m->hminff_kappa, ti, lc[RHG_LC_HMIN_FF_LIDX]);
      chi = (c->nH0[k] * 1.0E-29) * pe * kappa;
    } else {
      double Ck = (KBOLTZMANN * THETA0 * 1.0E-32), theta_n = 1.0,
             sqrt_theta = sqrt(THETA0 / T);
      int n;
      chi = 0.0;
      for (n = 1; n < 6; n++) {
        theta_n *= sqrt_theta;
        chi += theta_n * lc[RHG_LC_HFF_LONG0 + n - 1];
      }
      chi *= c->nH0[k] * (ne * Ck);
    }
    chi_ai[k] += chi;
    eta_ai[k] += chi * Bnu;
    /* opacity fudge factor of H-, rh/background.c:362-369 (the metal one,
       background.c:438-451, is in the metal continua's cross-sections) */
    chi_ai[k] *= lc[RHG_LC_HMIN_FUDGE];
    eta_ai[k] *= lc[RHG_LC_HMIN_FUDGE];
    /* OH and CH bound-free, rh/ohchbf.c:33-394, 402-690 */
    if (lc[RHG_LC_OH_EIDX] >= 0.0) {
      if (T < m->oh_T[0] || T > m->oh_T[NTOH - 1]) { chi = 0.0; eta = 0.0; }
      else {
        /* at T equal to the last table temperature the reference reads
           one element past the table, multiplied by (T - T[i2]) = 0 */
        int i2 = locate_asc(NTOH, m->oh_T, T);
        double ti = (i2 == NTOH - 1) ? (double) i2 :
          (double) i2 + (T - m->oh_T[i2]) / (m->oh_T[i2 + 1] - m->oh_T[i2]);
        double kappa = exp(LG10 * bilinear(NTOH, NEOH, m->oh_cross, ti,
                                           lc[RHG_LC_OH_EIDX])) * SQ(CM_TO_M);
        stim = exp(-lc[RHG_LC_HC_KLA] / T);
        chi = c->nOH[k] * (1.0 - stim) * kappa;
        eta = c->nOH[k] * lc[RHG_LC_TWOHNU3_C2] * stim * kappa;
      }
      chi_ai[k] += chi;
      eta_ai[k] += eta;
    }
    if (lc[RHG_LC_CH_EIDX] >= 0.0) {
      if (T < m->ch_T[0] || T > m->ch_T[NTCH - 1]) { chi = 0.0; eta = 0.0; }
      else {
        /* at T equal to the last table temperature the reference reads
           one element past the table, multiplied by (T - T[i2]) = 0 */
        int i2 = locate_asc(NTCH, m->ch_T, T);
        double ti = (i2 == NTCH - 1) ? (double) i2 :
          (double) i2 + (T - m->ch_T[i2]) / (m->ch_T[i2 + 1] - m->ch_T[i2]);
        double kappa = exp(LG10 * bilinear(NTCH, NECH, m->ch_cross, ti,
                                           lc[RHG_LC_CH_EIDX])) * SQ(CM_TO_M);
        stim = exp(-lc[RHG_LC_HC_KLA] / T);
        chi = c->nCH[k] * (1.0 - stim) * kappa;
        eta = c->nCH[k] * lc[RHG_LC_TWOHNU3_C2] * stim * kappa;
      }
      chi_ai[k] += chi;
      eta_ai[k] += eta;
    }
    /* H bound-free, rh/hydrogen.c:137-208 */
    if (hasH) {
      double expla;
      chi = eta = 0.0;
      for (e = cp0; e < cp1; e++) {
        int ci = m->lam_cont_idx[e];
        if (m->cont_atom[ci] != 0) continue;
        {
          double sigma = m->lam_cont_alpha[e], gijk;
          int li = m->atom_level_ptr[0] + m->cont_i[ci];
          expla = exp(-lc[RHG_LC_BF_HC_KLA] / T);
          gijk = ns[li] / c->npstar[k] * expla;
          chi += sigma * (1.0 - expla) * ns[li];
          eta += lc[RHG_LC_BF_TWOHNU3_C2] * gijk * sigma * c->np[k];
        }
      }
      chi_ai[k] += chi;
      eta_ai[k] += eta;
    }
    /* H free-free, rh/hydrogen.c:213-243, Gaunt factor :264-281 */
    {
      double y = (lc[RHG_LC_GFF_CY] * T) / (HPLANCK * CLIGHT), x3 = lc[RHG_LC_GFF_X3], g_ff;
      double gIII = 1.0 + 0.1728 * x3 * (1.0 + y) -
                    0.0496 * SQ(x3) * (1.0 + (1.0 + y) * 0.33333333 * y);
      g_ff = (gIII > 1.0) ? gIII : 1.0;
      stim = 1.0 - exp(-lc[RHG_LC_HC_KLA] / T);
      chi = sigma_ff / sqrt(T) * lc[RHG_LC_HFF_NU3] * ne * c->np[k] * stim * g_ff;
      chi_ai[k] += chi;
      eta_ai[k] += chi * Bnu;
    }
    /* H2+ free-free, rh/hydrogen.c:797-932 */
    if (lc[RHG_LC_H2P_LIDX] >= 0.0) {
      double ti = clamp_index(NTEMP_H2P, m->h2pff_temp, T);
      double kappa = bilinear(NTEMP_H2P, NFF_H2P, m->h2pff_kappa, ti, lc[RHG_LC_H2P_LIDX]);
      chi = (c->nH0[k] * 1.0E-29) * (c->np[k] * 1.0E-20) * kappa;
      chi_ai[k] += chi;
      eta_ai[k] += chi * Bnu;
    }
    /* H2- free-free, rh/hydrogen.c:620-793 */
    if (lc[RHG_LC_H2M_LIDX] >= 0.0) {
      if (c->nH2[k] > 0.0) {
        double ti = clamp_index(NTHETA_H2, m->h2mff_theta, THETA0 / T);
        double pe = ne * KBOLTZMANN * T;
        double kappa = bilinear(NTHETA_H2, NFF_H2, m->h2mff_kappa, ti, lc[RHG_LC_H2M_LIDX]);
        chi = (c->nH2[k] * 1.0E-29) * pe * kappa;
      } else chi = 0.0;
      chi_ai[k] += chi;
      eta_ai[k] += chi * Bnu;
    }
    /* metal bound-free, rh/metal.c:69-172 */
    {
      double expla = 0.0;
      int have = 0;
      chi = eta = 0.0;
      for (e = cp0; e < cp1; e++) {
        int ci = m->lam_cont_idx[e], ia = m->cont_atom[ci];
        int l0, li, lj;
        double gijk, alpha_la;
        if (ia == 0) continue;
        if (!have) { expla = exp(-lc[RHG_LC_BF_HC_KLA] / T); have = 1; }
        l0 = m->atom_level_ptr[ia];
        li = l0 + m->cont_i[ci];
        lj = l0 + m->cont_j[ci];
        alpha_la = m->lam_cont_alpha[e];
        gijk = ns[li] / ns[lj] * expla;
        chi += alpha_la * (1.0 - expla) * ns[li];
        eta += lc[RHG_LC_BF_TWOHNU3_C2] * gijk * alpha_la * ns[lj];
      }
      chi_ai[k] += chi * 1.0;
      eta_ai[k] += eta * 1.0;
    }
    (void) x;
  }
}

/* Kurucz-line opacity for the ray leaving the top (to_obs = 1),
   rh/kurucz.c:511-827.  Writes chi/eta [4][Ndep] from zero; returns 0 if
   the wavelength has no contributing line. */
/* Kurucz-line opacity at wavelength ls for the ray of direction cosine mu
   (mu < 0: the inward ray's Doppler shift, rh/kurucz.c:744-749);
   sel >= 0 keeps only model line sel (its chi_l phi and eta_l phi, as
   rh/kurucz.c:696-699 records for the log gf response). */
static int line_opacity_sel(const rhg_model_t *m, const column_t *c, int ls,
                            double mu, double *chi, double *eta, double *chip, int sel)
{
  const double lambda = m->lamconst[(size_t) ls * RHG_LC_NCONST + RHG_LC_LAMBDA];
  int p0 = m->lam_line_ptr[ls], p1 = m->lam_line_ptr[ls + 1], e, k, N = c->Ndep;
  const int nlev = m->atom_level_ptr[m->natom];
  if (p1 == p0) return 0;
  for (k = 0; k < 4 * N; k++) { chi[k] = 0.0; eta[k] = 0.0; }
  if (chip) for (k = 0; k < 3 * N; k++) chip[k] = 0.0;
  for (e = p0; e < p1; e++) {
    int l = m->lam_line_idx[e], ie = m->line_el[l], st = m->line_stage[l];
    if (sel >= 0 && l != sel) continue;
    double lambda0 = m->line_lambda0[l], weight = m->el_weight[ie];
    double vtherm = 2.0 * KBOLTZMANN / (AMU * weight);
    const double *lp = m->el_logpf + m->el_pf_ptr[ie] + (size_t) st * m->npf;
    int c0 = m->line_comp_ptr[l], c1 = m->line_comp_ptr[l + 1];
    for (k = 0; k < N; k++) {
      double T = c->T[k], vbroad, v, sv, phi, phi_Q = 0, phi_U = 0, phi_V = 0;
      double psi_Q = 0, psi_U = 0, psi_V = 0;
      double pf = linear1(m->npf, m->Tpf, lp, T);
      vbroad = sqrt(vtherm * T + SQ(c->vturb[k]));
      v = (lambda / lambda0 - 1.0) * CLIGHT / vbroad;
      if (c->moving) v += (mu * c->vz[k]) / vbroad;
      sv = 1.0 / (SQRTPI * vbroad);
      if (m->line_Grad[l] != 0.0) {
        double GvdW, adamp;
        switch (m->line_vdw[l]) {
        case RHG_VDW_UNSOLD: GvdW = m->line_cross[l] * pow(T, 0.3); break;
        case RHG_VDW_BARKLEM:
          GvdW = m->line_cross[l] * pow(T, (1.0 - m->line_alpha[l]) / 2.0); break;
        default: GvdW = m->line_GvdW[l]; break;
        }
        adamp = (m->line_Grad[l] + m->line_GStark[l] * c->ne[k] +
                 GvdW * (c->nHtot[k] - c->np[k])) *
                (lambda0 * NM_TO_M) / (4.0 * PI * vbroad);
        if (m->line_polar[l]) {
          double sin2_gamma = 1.0 - SQ(c->cosg[k]);
          double vB = m->line_larmor[l] * c->B[k] / vbroad, sign = 1.0;
          double phi_sm = 0, phi_pi = 0, phi_sp = 0, phi_sigma, phi_delta, H, F;
          double psi_sm = 0, psi_pi = 0, psi_sp = 0, psi_sigma, psi_delta;
          int nz;
          for (nz = c0; nz < c1; nz++) {
            H = rho_voigt_humlicek(adamp, v - m->comp_shift[nz] * vB, &F);
            switch (m->comp_q[nz]) {
            case -1: phi_sm += m->comp_strength[nz] * H; psi_sm += m->comp_strength[nz] * F; break;
            case 0: phi_pi += m->comp_strength[nz] * H; psi_pi += m->comp_strength[nz] * F; break;
            case 1: phi_sp += m->comp_strength[nz] * H; psi_sp += m->comp_strength[nz] * F; break;
            }
          }
          phi_sigma = phi_sp + phi_sm;
          phi_delta = 0.5 * phi_pi - 0.25 * phi_sigma;
          phi = (phi_delta * sin2_gamma + 0.5 * phi_sigma) * sv;
          phi_Q = sign * phi_delta * sin2_gamma * c->cos2x[k] * sv;
          phi_U = phi_delta * sin2_gamma * c->sin2x[k] * sv;
          phi_V = sign * 0.5 * (phi_sp - phi_sm) * c->cosg[k] * sv;
          /* the Faraday-Voigt profiles (rh/kurucz.c:811-819) */
          if (chip) {
            psi_sigma = psi_sp + psi_sm;
            psi_delta = 0.5 * psi_pi - 0.25 * psi_sigma;
            psi_Q = sign * psi_delta * sin2_gamma * c->cos2x[k] * sv;
            psi_U = psi_delta * sin2_gamma * c->sin2x[k] * sv;
            psi_V = sign * 0.5 * (psi_sp - psi_sm) * c->cosg[k] * sv;
          }
        } else {
          phi = rho_voigt_armstrong(adamp, v) * sv;
        }
      } else {
        phi = (fabs(v) <= MAX_GAUSS_DOPPLER) ? exp(-v * v) : 0.0;
        phi = phi * sv;
      }
      if (phi) {
        double kT = 1.0 / (KBOLTZMANN * T);
        double ni_gi = c->nel[((size_t) k * m->nel + ie) * NSTAGE_MAX + st] *
                       exp(-m->line_Ei[l] * kT - pf);
        double nj_gj = ni_gi * exp(-m->line_hc_la[l] * kT);
        double chi_l = m->line_Bijhc4pi[l] * (ni_gi - nj_gj);
        double eta_l = m->line_Bijhc4pi[l] * m->line_twohnu3_c2[l] * nj_gj;
        chi[k] += chi_l * phi;
        eta[k] += eta_l * phi;
        if (m->line_polar[l] && m->line_Grad[l]) {
          chi[N + k] += chi_l * phi_Q;
          chi[2 * N + k] += chi_l * phi_U;
          chi[3 * N + k] += chi_l * phi_V;
          eta[N + k] += eta_l * phi_Q;
          eta[2 * N + k] += eta_l * phi_U;
          eta[3 * N + k] += eta_l * phi_V;
          if (chip) {   /* rh/kurucz.c:711-715 */
            chip[k] += chi_l * psi_Q;
            chip[N + k] += chi_l * psi_U;
            chip[2 * N + k] += chi_l * psi_V;
          }
        }
      }
    }
  }
  (void) nlev;
  return 1;
}

static int line_opacity(const rhg_model_t *m, const column_t *c, int ls,
                        double mu, double *chi, double *eta, double *chip)
{
  return line_opacity_sel(m, c, ls, mu, chi, eta, chip, -1);
}

struct nlte_out { double *pops; int *niter; double *dpops; int ndep_max; double *stokes; double mu; };
static int nlte_column(const rhg_model_t *m, const rhg_nlte_t *na, const column_t *c,
                       struct nlte_out *o);

/* the depth scales convertScales leaves (rho_scales): where to put tau_ref
   and cmass of the column being solved on this thread (NULL: nowhere) */
static _Thread_local double *g_tau_out = NULL, *g_cmass_out = NULL;

static int synth_column_ex(const rhg_model_t *m, const rhg_columns_t *cols, int ic,
                           int fp32, double *stokes, double *height_out,
                           double *chiref_out, int diag_ls, double *diag_chi,
                           double *diag_eta, double *diag_pop, int npar,
                           const int *par_line, double *rf, const rhg_nlte_t *na,
                           struct nlte_out *no)
{
  const int N = cols->ndep[ic], Nmax = cols->ndep_max, nspect = m->nspect;
  const int nlev = m->atom_level_ptr[m->natom], natom = m->natom;
  const double *atm = cols->atmos + (size_t) ic * 9 * Nmax;
  const double mu = cols->mu;
  column_t c;
  double *mem, *chi_ai, *eta_ai, *chi_c, *eta_c, *lchi, *leta, *chiref, *S4, *I4,
         *molbuf, *scr, *chipb;
  int k, ls, i, nout = 0;

  memset(&c, 0, sizeof(c));
  c.Ndep = N;
  mem = calloc((size_t) N * (24 + nlev + natom + m->nel * NSTAGE_MAX + 30) +
               (size_t) 4 * N * 8 + (size_t) 3 * N + 64, sizeof(double));
  if (!mem) return -1;
  scr = mem;
#define TAKE(p, cnt) do { p = scr; scr += (cnt); } while (0)
  TAKE(c.T, N); TAKE(c.ne, N); TAKE(c.vz, N); TAKE(c.vturb, N); TAKE(c.B, N);
  TAKE(c.nHtot, N); TAKE(c.cosg, N); TAKE(c.cos2x, N); TAKE(c.sin2x, N);
  TAKE(c.tau_ref, N); TAKE(c.cmass, N); TAKE(c.height, N);
  TAKE(c.nstar, (size_t) N * nlev); TAKE(c.ntotal, (size_t) N * natom);
  TAKE(c.nHmin, N); TAKE(c.nH2, N); TAKE(c.nOH, N); TAKE(c.nCH, N); TAKE(c.nH0, N);
  TAKE(c.np, N); TAKE(c.npstar, N);
  TAKE(c.nel, (size_t) N * m->nel * NSTAGE_MAX);
  TAKE(chi_ai, N); TAKE(eta_ai, N); TAKE(chi_c, 4 * N); TAKE(eta_c, 4 * N);
  TAKE(lchi, 4 * N); TAKE(leta, 4 * N); TAKE(chiref, N); TAKE(S4, 4 * N);
  TAKE(I4, 4 * N); TAKE(molbuf, 32); TAKE(chipb, 3 * N);
#undef TAKE

  /* units and scales, rh/rhf1d/pyrh_compute1dray.c:230-278 */
  for (k = 0; k < N; k++) {
    double sc = atm[k];
    c.T[k] = atm[Nmax + k];
    c.ne[k] = atm[2 * Nmax + k] / CUBE(CM_TO_M);
    c.vz[k] = atm[3 * Nmax + k] * KM_TO_M;
    c.vturb[k] = atm[4 * Nmax + k] * KM_TO_M;
    c.B[k] = atm[5 * Nmax + k] / 1e4;
    c.nHtot[k] = atm[8 * Nmax + k] / CUBE(CM_TO_M);
    if (cols->atm_scale == 0) c.tau_ref[k] = POW10(sc);
    else if (cols->atm_scale == 1) c.cmass[k] = POW10(sc) * (G_TO_KG / SQ(CM_TO_M));
    else c.height[k] = sc * KM_TO_M;
    /* Bproject, rh/rhf1d/project.c:38-80 */
    {
      double gam = atm[6 * Nmax + k], chB = atm[7 * Nmax + k];
      if (mu == 1.0) {
        c.cosg[k] = cos(gam);
        c.cos2x[k] = cos(2.0 * chB);
        c.sin2x[k] = sin(2.0 * chB);
      } else {
        double mux = sqrt(1.0 - SQ(mu)), muy = 0.0, csc_theta = 1.0 / sqrt(1.0 - SQ(mu));
        double sin_gamma = sin(gam), bx = sin_gamma * cos(chB), by = sin_gamma * sin(chB),
               bz = cos(gam), b3 = mux * bx + muy * by + mu * bz,
               b1 = csc_theta * (bz - mu * b3), b2 = csc_theta * (muy * bx - mux * by);
        c.cosg[k] = b3;
        c.cos2x[k] = (SQ(b1) - SQ(b2)) / (1.0 - SQ(b3));
        c.sin2x[k] = 2.0 * b1 * b2 / (1.0 - SQ(b3));
      }
    }
  }
  c.moving = 0;
  for (k = 0; k < N; k++)
    if (fabs(c.vz[k]) >= m->vmacro_tresh) { c.moving = 1; break; }

  /* LTE populations and chemical equilibrium, rh/background.c:164-169 */
  for (k = 0; k < N; k++) {
    double *ns = c.nstar + (size_t) k * nlev, *nt = c.ntotal + (size_t) k * natom;
    int ia;
    for (ia = 0; ia < natom; ia++) {
      nt[ia] = m->atom_abund[ia] * c.nHtot[k];
      ltepops(m, ia, c.T[k], c.ne[k], nt[ia], ns + m->atom_level_ptr[ia]);
    }
    c.nHmin[k] = chemeq(m, c.T[k], c.ne[k], c.nHtot[k], ns, nt, molbuf);
    c.nH2[k] = m->mol_h2 >= 0 ? molbuf[m->mol_h2] : 0.0;
    c.nOH[k] = m->mol_oh >= 0 ? molbuf[m->mol_oh] : 0.0;
    c.nCH[k] = m->mol_ch >= 0 ? molbuf[m->mol_ch] : 0.0;
    c.nH0[k] = ns[m->atom_level_ptr[0]];
    c.np[k] = ns[m->atom_level_ptr[1] - 1];
    c.npstar[k] = c.np[k];
    for (i = 0; i < m->nel; i++)
      ltepops_elem(m, i, c.T[k], c.ne[k], c.nHtot[k],
                   c.nel + ((size_t) k * m->nel + i) * NSTAGE_MAX);
  }

  /* background at the reference wavelength, then convertScales
     (rh/rhf1d/multiatmos.c:100-177) */
  background_continuum(m, &c, m->ref_index, chiref, eta_ai);
  if (na) {
    /* solving NLTE, the background opacity includes the scattering
       opacity (rh/background.c:461-463), also at the reference wavelength
       that sets the depth scale */
    const int lr = m->ref_index, nlev_all = m->atom_level_ptr[m->natom];
    for (k = 0; k < N; k++) {
      double sca = c.ne[k] * na->sigma_thomson;
      sca += na->ray_H[lr] * c.nH0[k];
      if (na->he_atom >= 0)
        sca += na->ray_He[lr] * c.nstar[(size_t) k * nlev_all + m->atom_level_ptr[na->he_atom]];
      sca += na->ray_H2[lr] * c.nH2[k];
      sca *= m->lamconst[(size_t) lr * RHG_LC_NCONST + RHG_LC_SCATT_FUDGE];
      chiref[k] += sca;
    }
  }
  if (line_opacity(m, &c, m->ref_index, mu, lchi, leta, NULL))
    for (k = 0; k < N; k++) chiref[k] += lchi[k];
  {
    double *rho = S4, h_zero, unity = 1.0;   /* S4 as scratch */
    for (k = 0; k < N; k++) rho[k] = (AMU * m->wght_per_H) * c.nHtot[k];
    if (cols->atm_scale == 0) {
      c.height[0] = 0.0;
      c.cmass[0] = (c.tau_ref[0] / chiref[0]) * rho[0];
      for (k = 1; k < N; k++) {
        c.height[k] = c.height[k - 1] - 2.0 * (c.tau_ref[k] - c.tau_ref[k - 1]) /
                      (chiref[k - 1] + chiref[k]);
        c.cmass[k] = c.cmass[k - 1] + 0.5 * (rho[k - 1] + rho[k]) *
                     (c.height[k - 1] - c.height[k]);
      }
    } else if (cols->atm_scale == 1) {
      c.height[0] = 0.0;
      c.tau_ref[0] = chiref[0] / rho[0] * c.cmass[0];
      for (k = 1; k < N; k++) {
        c.height[k] = c.height[k - 1] - 2.0 * (c.cmass[k] - c.cmass[k - 1]) /
                      (rho[k - 1] + rho[k]);
        c.tau_ref[k] = c.tau_ref[k - 1] + 0.5 * (chiref[k - 1] + chiref[k]) *
                       (c.height[k - 1] - c.height[k]);
      }
    }
    if (cols->atm_scale == 0 || cols->atm_scale == 1) {
      h_zero = linear1(N, c.tau_ref, c.height, unity);
      for (k = 0; k < N; k++) c.height[k] = c.height[k] - h_zero;
    }
    if (cols->atm_scale == 2 && N > 1) {
      c.cmass[0] = (c.nHtot[0] * m->totalAbund + c.ne[0]) * (KBOLTZMANN * c.T[0] / m->gravity);
      c.tau_ref[0] = 0.5 * chiref[0] * (c.height[0] - c.height[1]);
      if (c.tau_ref[0] > 1.0) c.tau_ref[0] = 0.0;
      for (k = 1; k < N; k++) {
        c.cmass[k] = c.cmass[k - 1] + 0.5 * (rho[k - 1] + rho[k]) * (c.height[k - 1] - c.height[k]);
        c.tau_ref[k] = c.tau_ref[k - 1] + 0.5 * (chiref[k - 1] + chiref[k]) *
                       (c.height[k - 1] - c.height[k]);
      }
    }
    if (g_tau_out) for (k = 0; k < N; k++) g_tau_out[k] = c.tau_ref[k];
    if (g_cmass_out) for (k = 0; k < N; k++) g_cmass_out[k] = c.cmass[k];
  }
  if (height_out) for (k = 0; k < N; k++) height_out[k] = c.height[k];
  if (chiref_out) for (k = 0; k < N; k++) chiref_out[k] = chiref[k];
  if (diag_pop) {
    for (k = 0; k < N; k++) {
      diag_pop[0 * N + k] = c.nHmin[k];
      diag_pop[1 * N + k] = c.nH2[k];
      diag_pop[2 * N + k] = c.nOH[k];
      diag_pop[3 * N + k] = c.nCH[k];
      diag_pop[4 * N + k] = c.nH0[k];
      diag_pop[5 * N + k] = c.np[k];
    }
  }
  if (diag_ls >= 0) {
    background_continuum(m, &c, diag_ls, chi_ai, eta_ai);
    for (k = 0; k < N; k++) { diag_chi[k] = chi_ai[k]; diag_eta[k] = eta_ai[k]; }
    for (k = N; k < 4 * N; k++) { diag_chi[k] = 0.0; diag_eta[k] = 0.0; }
    if (line_opacity(m, &c, diag_ls, mu, lchi, leta, NULL))
      for (k = 0; k < 4 * N; k++) { diag_chi[4 * N + k] = lchi[k]; diag_eta[4 * N + k] = leta[k]; }
    else
      for (k = 0; k < 4 * N; k++) { diag_chi[4 * N + k] = 0.0; diag_eta[4 * N + k] = 0.0; }
    free(mem);
    return 0;
  }

  if (na) {
    int rc = nlte_column(m, na, &c, no);
    free(mem);
    return rc;
  }

  /* formal solution of the emergent ray per wavelength (formal.c:157-275) */
  for (ls = 0; ls < nspect; ls++) {
    int has, solver, n;
    double B2[2], Iout[4] = {0, 0, 0, 0};
    if (ls == m->ref_index) continue;
    background_continuum(m, &c, ls, chi_ai, eta_ai);
    has = line_opacity(m, &c, ls, mu, lchi, leta, m->magneto_optical ? chipb : NULL);
    for (k = 0; k < N; k++) { chi_c[k] = chi_ai[k]; eta_c[k] = eta_ai[k]; }
    for (k = N; k < 4 * N; k++) { chi_c[k] = 0.0; eta_c[k] = 0.0; }
    if (has) for (k = 0; k < 4 * N; k++) { chi_c[k] += lchi[k]; eta_c[k] += leta[k]; }
    /* a BEZIER wavelength is angle-independent in a static column and
       takes Feautrier (rh/rhf1d/formal.c:98-101, 287-310); a polarizable
       background line keeps it angle-dependent (BEZIER_ANGLE) */
    solver = m->solver[ls];
    if (solver == RHG_SOLVER_BEZIER && !c.moving) solver = RHG_SOLVER_FEAUTRIER;
    if (solver == RHG_SOLVER_BEZIER_ANGLE) solver = RHG_SOLVER_BEZIER;
    B2[0] = rho_planck(c.T[N - 2], m->lamconst[(size_t) ls * RHG_LC_NCONST]);
    B2[1] = rho_planck(c.T[N - 1], m->lamconst[(size_t) ls * RHG_LC_NCONST]);
    if (solver == RHG_SOLVER_STOKES) {
      for (n = 0; n < 4; n++)
        for (k = 0; k < N; k++) S4[n * N + k] = (0.0 + eta_c[n * N + k]) / chi_c[k];
      rho_formal_stokes(N, c.height, mu, chi_c, S4, chi_c + N,
                        m->magneto_optical ? chipb : NULL, B2, fp32, I4);
      for (n = 0; n < 4; n++) Iout[n] = I4[n * N];
    } else if (solver == RHG_SOLVER_BEZIER && rf) {
      /* the inward ray first, then the emergent one with the log gf
         response (rh/rhf1d/formal.c:168-172, 281-284) */
      double *dchi = calloc((size_t) 2 * npar * N + 1, sizeof(double)), *deta;
      int p;
      if (!dchi) { free(mem); return -1; }
      deta = dchi + (size_t) npar * N;
      /* pyrh keeps one background array per wavelength
         (rh/readj.c:241-268, filled by writeBackground rh/readj.c:300-345);
         the last pass that wrote it is the emergent ray's
         (rh/background.c:470-580), so the inward ray runs on the emergent
         ray's opacity and source function */
      for (k = 0; k < N; k++) S4[k] = (0.0 + eta_c[k]) / chi_c[k];
      rho_formal_bezier_in(N, c.height, mu, chi_c, S4,
                           m->lamconst[(size_t) ls * RHG_LC_NCONST + RHG_LC_ITOP], I4);
      for (p = 0; p < npar; p++) {
        if (par_line[p] < 0) continue;
        if (!line_opacity_sel(m, &c, ls, mu, lchi, leta, NULL, par_line[p])) continue;
        for (k = 0; k < N; k++) {
          dchi[(size_t) p * N + k] = lchi[k] * LG10;
          deta[(size_t) p * N + k] = leta[k] * LG10;
        }
      }
      rho_formal_bezier_rf(N, c.height, mu, chi_c, S4, B2, npar, dchi, deta, I4,
                           rf + (size_t) nout * npar);
      Iout[0] = I4[0];
      free(dchi);
    } else if (solver == RHG_SOLVER_BEZIER) {
      for (k = 0; k < N; k++) S4[k] = (0.0 + eta_c[k]) / chi_c[k];
      rho_formal_bezier(N, c.height, mu, chi_c, S4, B2, I4);
      Iout[0] = I4[0];
    } else {
      for (k = 0; k < N; k++) S4[k] = (0.0 + eta_c[k] + 0.0) / chi_c[k];
      Iout[0] = formal_feautrier_h0(N, c.height, mu, chi_c, S4, B2,
                                    m->lamconst[(size_t) ls * RHG_LC_NCONST + RHG_LC_ITOP]);
    }
    for (n = 0; n < 4; n++) stokes[n * (nspect - 1) + nout] = Iout[n];
    nout++;
  }
  free(mem);
  return 0;
}

static int synth_column(const rhg_model_t *m, const rhg_columns_t *cols, int ic,
                        int fp32, double *stokes, double *height_out,
                        double *chiref_out)
{
  return synth_column_ex(m, cols, ic, fp32, stokes, height_out, chiref_out,
                         -1, NULL, NULL, NULL, 0, NULL, NULL, NULL, NULL);
}

/* Diagnostic: continuum (chi[0:N], eta[0:N]) and line (chi/eta[4N:8N])
   opacity of column ic at wavelength ls for the emergent ray, and the
   populations nHmin, nH2, nOH, nCH, nH(1), np ([6][N]). */
int rho_background(const rhg_model_t *m, const rhg_columns_t *cols, int ic, int ls,
                   double *chi8, double *eta8, double *pop6)
{
  return synth_column_ex(m, cols, ic, 0, NULL, NULL, NULL, ls, chi8, eta8, pop6, 0, NULL,
                         NULL, NULL, NULL);
}

int rho_synth_lte(const rhg_model_t *m, const rhg_columns_t *cols,
                  int inverse_fp32, int nthreads, double *stokes_out,
                  double *height_out, double *chiref_out)
{
  int ic, err = 0;
  const int nout = m->nspect - 1;
  if (m->nnuclei + m->nmol > NEQ_MAX) return RHG_EUNSUPPORTED;
  for (ic = 0; ic < m->nel; ic++) if (m->el_nstage[ic] > NSTAGE_MAX) return RHG_EUNSUPPORTED;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(|:err)
#endif
  for (ic = 0; ic < cols->ncol; ic++) {
    if (synth_column(m, cols, ic, inverse_fp32, stokes_out + (size_t) ic * 4 * nout,
                     height_out ? height_out + (size_t) ic * cols->ndep_max : NULL,
                     chiref_out ? chiref_out + (size_t) ic * cols->ndep_max : NULL))
      err |= 1;
  }
  (void) nthreads;
  return err ? RHG_ENOMEM : RHG_OK;
}

/* The reference's analytic d I / d log gf with the synthesis
   (compute1d(get_atomic_rfs=True), rh/rhf1d/bezier_1D.c:416-536): rf_out is
   [ncol][nspect-1][npar], zero at the wavelengths the scalar Bezier does
   not solve. */
int rho_rf_loggf(const rhg_model_t *m, const rhg_columns_t *cols, int npar,
                 const int *par_line, int nthreads, double *stokes_out, double *rf_out)
{
  int ic, err = 0;
  const int nout = m->nspect - 1;
  if (m->nnuclei + m->nmol > NEQ_MAX) return RHG_EUNSUPPORTED;
  memset(rf_out, 0, (size_t) cols->ncol * nout * npar * sizeof(double));
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(|:err)
#endif
  for (ic = 0; ic < cols->ncol; ic++) {
    if (synth_column_ex(m, cols, ic, 0, stokes_out + (size_t) ic * 4 * nout, NULL, NULL, -1,
                        NULL, NULL, NULL, npar, par_line,
                        rf_out + (size_t) ic * nout * npar, NULL, NULL))
      err |= 1;
  }
  (void) nthreads;
  return err ? RHG_ENOMEM : RHG_OK;
}


/* ------------------------------------------------------------------ */
/* NLTE: one active atom, accelerated lambda iteration as rhf1d() runs */
/* it (rh/rhf1d/pyrh_compute1dray.c:280-352, rh/iterate.c:48-143).     */

static double m_lambda(const rhg_model_t *m, int ls)
{
  return m->lamconst[(size_t) ls * RHG_LC_NCONST + RHG_LC_LAMBDA];
}

/* RH's splineEval with the table's second derivatives M precomputed
   (rh/spline.c:70-97) */
static double spline_eval_M(int n, const double *x, const double *y, const double *M,
                            double xv)
{
  int j;
  double hj, fx, fx1;
  if (xv <= x[0]) return y[0];
  if (xv >= x[n - 1]) return y[n - 1];
  j = locate_asc(n, x, xv);
  hj = x[j + 1] - x[j];
  fx = (xv - x[j]) / hj;
  fx1 = 1 - fx;
  return fx1 * y[j] + fx * y[j + 1] +
         (fx1 * (SQ(fx1) - 1) * M[j] + fx * (SQ(fx) - 1) * M[j + 1]) * SQ(hj) / 6.0;
}

/* Scalar cubic Bezier along one ray with the diagonal operator
   (rh/rhf1d/bezier_1D.c:306-541): to_obs = 1 from the thermalised bottom
   (B2 = B(T[N-2]), B(T[N-1])), to_obs = 0 from a dark top.  I and Psi at
   every depth. */
static void bezier_ray(int Ndep, const double *z, double zmu, const double *chi,
                       const double *S, const double *B2, int to_obs, double *I, double *Psi)
{
  int dk, k_start, k_end, k;
  double dtau_uw, dtau_dw = 0.0, dS_uw, I_upw, c1, c2, w[3], dsup, dsdn, dt03, co[5], dS_up,
         dS_c, dchi_up, dchi_c, dchi_dn = 0.0, dsdn2;
  if (to_obs) { dk = -1; k_start = Ndep - 1; k_end = 0; }
  else        { dk = 1;  k_start = 0;        k_end = Ndep - 1; }
  dtau_uw = 0.5 * zmu * (chi[k_start] + chi[k_start + dk]) *
            fabs(z[k_start] - z[k_start + dk]);
  I_upw = to_obs ? B2[1] - (B2[0] - B2[1]) / dtau_uw : 0.0;
  I[k_start] = I_upw;
  Psi[k_start] = 0.0;
  k = k_start + dk;
  dsup = fabs(z[k] - z[k - dk]) * zmu;
  dsdn = fabs(z[k + dk] - z[k]) * zmu;
  dchi_up = (chi[k] - chi[k - dk]) / dsup;
  dchi_c = cent_deriv(dsup, dsdn, chi[k - dk], chi[k], chi[k + dk]);
  c1 = MAXV(chi[k] - (dsup / 3.0) * dchi_c, 0.0);
  c2 = MAXV(chi[k - dk] + (dsup / 3.0) * dchi_up, 0.0);
  dtau_uw = dsup * (chi[k] + chi[k - dk] + c1 + c2) * 0.25;
  dS_up = (S[k] - S[k - dk]) / dtau_uw;
  for (k = k_start + dk; k != k_end + dk; k += dk) {
    if (k != k_end) {
      dsdn = fabs(z[k + dk] - z[k]) * zmu;
      if (abs(k - k_end) > 1) {
        dsdn2 = fabs(z[k + 2 * dk] - z[k + dk]) * zmu;
        dchi_dn = cent_deriv(dsdn, dsdn2, chi[k], chi[k + dk], chi[k + 2 * dk]);
      } else
        dchi_dn = (chi[k + dk] - chi[k]) / dsdn;
      c1 = MAXV(chi[k] + (dsdn / 3.0) * dchi_c, 0.0);
      c2 = MAXV(chi[k + dk] - (dsdn / 3.0) * dchi_dn, 0.0);
      dtau_dw = dsdn * (chi[k] + chi[k + dk] + c1 + c2) * 0.25;
      dt03 = dtau_uw / 3.0;
      rho_bezier3_coeffs(dtau_uw, co);
      dS_c = cent_deriv(dtau_uw, dtau_dw, S[k - dk], S[k], S[k + dk]);
      c1 = MAXV(S[k] - dt03 * dS_c, 0.0);
      c2 = MAXV(S[k - dk] + dt03 * dS_up, 0.0);
      I[k] = I_upw * co[4] + co[0] * S[k] + co[1] * S[k - dk] + co[2] * c1 + co[3] * c2;
      Psi[k] = co[0] + co[2];
    } else {
      dtau_uw = 0.5 * zmu * (chi[k] + chi[k - dk]) * fabs(z[k] - z[k - dk]);
      dS_uw = -(S[k] - S[k - dk]) / dtau_uw;
      rho_w3(dtau_uw, w);
      I[k] = (1.0 - w[0]) * I_upw + w[0] * S[k] + w[1] * dS_uw;
      Psi[k] = w[0] - w[1] / dtau_uw;
    }
    I_upw = I[k];
    dsup = dsdn;
    dchi_up = dchi_c;
    dchi_c = dchi_dn;
    dtau_uw = dtau_dw;
    dS_up = dS_c;
  }
  (void) dsup;
  (void) dchi_up;
}

/* Feautrier, STANDARD order, dark top, thermalised bottom, with P and the
   two-sided diagonal operator at every depth (rh/rhf1d/feautrier.c:56-202) */
static void feautrier_full(int Ndep, const double *z, double mu, const double *chi,
                           const double *S, const double *B2, double *P, double *Psi,
                           double *scr)
{
  double *dtau = scr, *abc = scr + Ndep, *A1 = scr + 2 * Ndep, *C1 = scr + 3 * Ndep,
         *F = scr + 4 * Ndep, *G = scr + 5 * Ndep, *ztmp = scr + 6 * Ndep, *Stmp = scr + 7 * Ndep;
  double zmu = 0.5 / mu, r0 = 0.0, h0 = 0.0, rN = 0.0, hN, f0, fN, dtau_mid;
  int k;
  for (k = 0; k < Ndep - 1; k++) dtau[k] = zmu * (chi[k] + chi[k + 1]) * (z[k] - z[k + 1]);
  f0 = (1.0 - r0) / (1.0 + r0);
  abc[0] = 1.0 + 2.0 * f0 / dtau[0];
  C1[0] = 2.0 / SQ(dtau[0]);
  Stmp[0] = S[0] + 2.0 * h0 / ((1.0 + r0) * dtau[0]);
  hN = B2[1] - (B2[0] - B2[1]) / dtau[Ndep - 2];
  fN = (1.0 - rN) / (1.0 + rN);
  abc[Ndep - 1] = 1.0 + 2.0 * fN / dtau[Ndep - 2];
  A1[Ndep - 1] = 2.0 / SQ(dtau[Ndep - 2]);
  Stmp[Ndep - 1] = S[Ndep - 1] + 2.0 * hN / ((1.0 + rN) * dtau[Ndep - 2]);
  for (k = 1; k < Ndep - 1; k++) {
    dtau_mid = 0.5 * (dtau[k] + dtau[k - 1]);
    A1[k] = 1.0 / (dtau_mid * dtau[k - 1]);
    C1[k] = 1.0 / (dtau_mid * dtau[k]);
    abc[k] = 1.0;
    Stmp[k] = S[k];
  }
  F[0] = abc[0] / C1[0];
  ztmp[0] = Stmp[0] / (abc[0] + C1[0]);
  for (k = 1; k < Ndep - 1; k++) {
    F[k] = (abc[k] + A1[k] * F[k - 1] / (1.0 + F[k - 1])) / C1[k];
    ztmp[k] = (Stmp[k] + A1[k] * ztmp[k - 1]) / (C1[k] * (1.0 + F[k]));
  }
  P[Ndep - 1] = (Stmp[Ndep - 1] + A1[Ndep - 1] * ztmp[Ndep - 2]) /
                (abc[Ndep - 1] + A1[Ndep - 1] * (F[Ndep - 2] / (1.0 + F[Ndep - 2])));
  for (k = Ndep - 2; k >= 0; k--) P[k] = P[k + 1] / (1.0 + F[k]) + ztmp[k];
  G[Ndep - 1] = abc[Ndep - 1] / A1[Ndep - 1];
  for (k = Ndep - 2; k >= 1; k--) G[k] = (abc[k] + C1[k] * G[k + 1] / (1.0 + G[k + 1])) / A1[k];
  Psi[0] = 1.0 / (abc[0] + C1[0] * G[1] / (1.0 + G[1]));
  for (k = 1; k < Ndep - 1; k++)
    Psi[k] = 1.0 / (abc[k] + A1[k] * F[k - 1] / (1.0 + F[k - 1]) + C1[k] * G[k + 1] / (1.0 + G[k + 1]));
  Psi[Ndep - 1] = 1.0 / (abc[Ndep - 1] + A1[Ndep - 1] * F[Ndep - 2] / (1.0 + F[Ndep - 2]));
}

#define NLEV_MAX 16

typedef struct {
  const rhg_model_t *m;
  const rhg_nlte_t *na;
  const column_t *c;
  int N, nl, nrays;
  double *n, *nstar, *ntotal, *vbroad, *C, *Gamma, *J, *Jdag;
  double *phi, *wphi;               /* per line: [la][mu][dir][k], [k] */
  double *adamp_l;                  /* per line: damping [k] */
  const int *phi_off;               /* line -> offset into phi */
  double *bg_chi, *bg_eta, *bg_sca; /* [nspect][N] (active wavelengths) */
  double *chi, *S, *I, *Psi, *eta_at, *Ieff, *scr;
  double *V, *gij, *wla;            /* [MAX][N] per transition of the set */
  double *chi_up, *chi_down, *Uji;  /* [nl][N] */
  FILE *dump;                       /* first formal pass only */
} nlte_ws;

static void nlte_dump(FILE *f, const char *tag, int idx, const double *x, int n);

/* Opacity of the active set at wavelength ls for ray (mu, to_obs):
   Vij, gij and wla per transition, the set's chi and the atom's eta
   (rh/opacity.c:62-300) */
static void nlte_opacity(nlte_ws *w, int ls, int mu, int to_obs, int initialize)
{
  const rhg_nlte_t *na = w->na;
  const double hc = HPLANCK * CLIGHT, fourPI = 4.0 * PI, hc_4PI = hc / fourPI,
               twohc = 2.0 * hc / CUBE(NM_TO_M), hc_k = hc / (KBOLTZMANN * NM_TO_M);
  const double *T = w->c->T;
  const int N = w->N, p0 = na->as_ptr[ls], p1 = na->as_ptr[ls + 1];
  int e, k;
  for (k = 0; k < N; k++) { w->chi[k] = 0.0; w->eta_at[k] = 0.0; }
  for (e = p0; e < p1; e++) {
    const int n = e - p0, t = na->as_tr[e], i = na->tr_i[t], j = na->tr_j[t];
    const int la = ls - na->tr_nblue[t];
    const double *ni = w->n + (size_t) i * N, *nj = w->n + (size_t) j * N;
    double *V = w->V + (size_t) n * N, *gij = w->gij + (size_t) n * N, *wla = w->wla + (size_t) n * N;
    double twohnu3_c2;
    if (na->tr_type[t] == 0) {
      const int l = na->tr_line[t];
      const double *phi = w->phi + w->phi_off[l] +
                          ((size_t) (la * w->nrays + mu) * 2 + to_obs) * N;
      const double gijk = na->ln_Bji[l] / na->ln_Bij[l];
      const double Bijxhc_4PI = hc_4PI * na->ln_Bij[l] * 1.0;
      twohnu3_c2 = na->ln_Aji[l] / na->ln_Bji[l];
      for (k = 0; k < N; k++) { gij[k] = gijk; V[k] = Bijxhc_4PI * phi[k]; }
      if (initialize) {
        const double wlambda = na->tr_wlam[na->tr_ptr[t] + la];
        for (k = 0; k < N; k++) wla[k] = wlambda * w->wphi[(size_t) l * N + k] / hc_4PI;
      }
    } else {
      const double lam = m_lambda(w->m, ls);
      twohnu3_c2 = twohc / CUBE(lam);
      if (initialize) {
        const double wlambda = na->tr_wlam[na->tr_ptr[t] + la];
        const double alpha = na->tr_alpha[na->tr_ptr[t] + la];
        for (k = 0; k < N; k++) {
          V[k] = alpha;
          gij[k] = w->nstar[(size_t) i * N + k] / w->nstar[(size_t) j * N + k] *
                   exp(-hc_k / (lam * T[k]));
          wla[k] = fourPI / HPLANCK * (wlambda / lam);
        }
      }
    }
    for (k = 0; k < N; k++) {
      w->chi[k] += V[k] * (ni[k] - gij[k] * nj[k]);
      w->eta_at[k] += twohnu3_c2 * gij[k] * V[k] * nj[k];
    }
  }
}

static double nlte_twohnu3(const nlte_ws *w, int ls, int t)
{
  const rhg_nlte_t *na = w->na;
  if (na->tr_type[t] == 0) return na->ln_Aji[na->tr_line[t]] / na->ln_Bji[na->tr_line[t]];
  {
    const double lam = m_lambda(w->m, ls);
    return 2.0 * HPLANCK * CLIGHT / CUBE(NM_TO_M) / CUBE(lam);
  }
}

/* addtoCoupling, rh/fillgamma.c:251-332 */
static void nlte_coupling(nlte_ws *w, int ls)
{
  const rhg_nlte_t *na = w->na;
  const int N = w->N, p0 = na->as_ptr[ls], p1 = na->as_ptr[ls + 1];
  int e, k;
  for (e = p0; e < p1; e++) {
    const int t = na->as_tr[e];
    for (k = 0; k < N; k++) {
      w->chi_up[(size_t) na->tr_i[t] * N + k] = 0.0;
      w->chi_down[(size_t) na->tr_j[t] * N + k] = 0.0;
      w->Uji[(size_t) na->tr_j[t] * N + k] = 0.0;
    }
  }
  for (e = p0; e < p1; e++) {
    const int n = e - p0, t = na->as_tr[e], i = na->tr_i[t], j = na->tr_j[t];
    const double twohnu3_c2 = nlte_twohnu3(w, ls, t);
    const double *V = w->V + (size_t) n * N, *gij = w->gij + (size_t) n * N,
                 *wla = w->wla + (size_t) n * N;
    const double *ni = w->n + (size_t) i * N, *nj = w->n + (size_t) j * N;
    if (twohnu3_c2) {
      for (k = 0; k < N; k++) {
        const double chicc = V[k] * wla[k] * (ni[k] - gij[k] * nj[k]);
        w->chi_up[(size_t) i * N + k] += chicc;
        w->chi_down[(size_t) j * N + k] += chicc;
        w->Uji[(size_t) j * N + k] += twohnu3_c2 * gij[k] * V[k];
      }
    }
  }
}

/* addtoGamma, rh/fillgamma.c:82-246 (Psi already divided by chi) */
static void nlte_gamma(nlte_ws *w, int ls, double wmu)
{
  const rhg_nlte_t *na = w->na;
  const int N = w->N, nl = w->nl, p0 = na->as_ptr[ls], p1 = na->as_ptr[ls + 1];
  int e, e2, k;
  for (k = 0; k < N; k++) w->Ieff[k] = w->I[k] - w->Psi[k] * w->eta_at[k];
  for (e = p0; e < p1; e++) {
    const int n = e - p0, t = na->as_tr[e], i = na->tr_i[t], j = na->tr_j[t];
    const double twohnu3_c2 = nlte_twohnu3(w, ls, t);
    const double *V = w->V + (size_t) n * N, *gij = w->gij + (size_t) n * N,
                 *wla = w->wla + (size_t) n * N;
    double *Gij = w->Gamma + (size_t) (i * nl + j) * N, *Gji = w->Gamma + (size_t) (j * nl + i) * N;
    for (k = 0; k < N; k++) {
      const double wlamu = V[k] * wla[k] * wmu;
      Gji[k] += w->Ieff[k] * wlamu;
      Gij[k] += (twohnu3_c2 + w->Ieff[k]) * gij[k] * wlamu;
    }
    for (k = 0; k < N; k++)
      Gij[k] -= w->chi_up[(size_t) i * N + k] * w->Psi[k] * w->Uji[(size_t) j * N + k] * wmu;
    for (e2 = p0; e2 < p1; e2++) {
      if (na->tr_j[na->as_tr[e2]] == i) {
        for (k = 0; k < N; k++)
          Gji[k] += w->chi_down[(size_t) j * N + k] * w->Psi[k] * w->Uji[(size_t) i * N + k] * wmu;
      }
    }
  }
}

/* Formal solution at one wavelength with the active set: rays, J and
   (eval_operator) the rate operator; returns max |1 - Jdag/J|
   (rh/rhf1d/formal.c:44-346) */
static double nlte_formal(nlte_ws *w, int ls, int eval_operator)
{
  const rhg_nlte_t *na = w->na;
  const column_t *c = w->c;
  const int N = w->N;
  const double lam = m_lambda(w->m, ls);
  double *Jl = w->J + (size_t) ls * N, B2[2], dJmax = 0.0;
  const double *chic = w->bg_chi + (size_t) ls * N, *etac = w->bg_eta + (size_t) ls * N,
               *scac = w->bg_sca + (size_t) ls * N;
  int k, mu, to_obs;
  for (k = 0; k < N; k++) { w->Jdag[k] = Jl[k]; Jl[k] = 0.0; }
  B2[0] = rho_planck(c->T[N - 2], lam);
  B2[1] = rho_planck(c->T[N - 1], lam);
  if (na->as_bezier[ls]) {
    for (mu = 0; mu < w->nrays; mu++) {
      const double wmu = 0.5 * na->wmu[mu];
      for (to_obs = 0; to_obs <= 1; to_obs++) {
        nlte_opacity(w, ls, mu, to_obs, mu == 0 && to_obs == 0);
        if (eval_operator) nlte_coupling(w, ls);
        for (k = 0; k < N; k++) {
          const double chi_as = w->chi[k];
          w->chi[k] = chi_as + chic[k];
          w->S[k] = w->eta_at[k] + etac[k];
          w->S[k] += scac[k] * w->Jdag[k];
        }
        for (k = 0; k < N; k++) w->S[k] /= w->chi[k];
        bezier_ray(N, c->height, 1.0 / na->muz[mu], w->chi, w->S, B2, to_obs, w->I, w->Psi);
        if (w->dump && to_obs) {
          char tg[24];
          snprintf(tg, sizeof(tg), "fs1_chi_%d", mu);
          nlte_dump(w->dump, tg, ls, w->chi, N);
          snprintf(tg, sizeof(tg), "fs1_S_%d", mu);
          nlte_dump(w->dump, tg, ls, w->S, N);
          snprintf(tg, sizeof(tg), "fs1_I_%d", mu);
          nlte_dump(w->dump, tg, ls, w->I, N);
          snprintf(tg, sizeof(tg), "fs1_eat_%d", mu);
          nlte_dump(w->dump, tg, ls, w->eta_at, N);
        }
        if (eval_operator) {
          for (k = 0; k < N; k++) w->Psi[k] /= w->chi[k];
          nlte_gamma(w, ls, wmu);
        }
        for (k = 0; k < N; k++) Jl[k] += wmu * w->I[k];
      }
    }
  } else {
    nlte_opacity(w, ls, 0, 0, 1);
    if (eval_operator) nlte_coupling(w, ls);
    for (k = 0; k < N; k++) {
      const double chi_as = w->chi[k];
      w->chi[k] = chi_as + chic[k];
      w->S[k] = (w->eta_at[k] + etac[k] + scac[k] * w->Jdag[k]) / w->chi[k];
    }
    for (mu = 0; mu < w->nrays; mu++) {
      feautrier_full(N, c->height, na->muz[mu], w->chi, w->S, B2, w->I, w->Psi, w->scr);
      if (w->dump) {
        char tg[24];
        snprintf(tg, sizeof(tg), "fe_chi_%d", mu);
        nlte_dump(w->dump, tg, ls, w->chi, N);
        snprintf(tg, sizeof(tg), "fe_S_%d", mu);
        nlte_dump(w->dump, tg, ls, w->S, N);
        snprintf(tg, sizeof(tg), "fe_P_%d", mu);
        nlte_dump(w->dump, tg, ls, w->I, N);
      }
      if (eval_operator) {
        for (k = 0; k < N; k++) w->Psi[k] /= w->chi[k];
        nlte_gamma(w, ls, na->wmu[mu]);
      }
      for (k = 0; k < N; k++) Jl[k] += w->I[k] * na->wmu[mu];
    }
  }
  for (k = 0; k < N; k++) {
    const double dJ = fabs(1.0 - w->Jdag[k] / Jl[k]);
    dJmax = MAXV(dJmax, dJ);
  }
  return dJmax;
}

/* Ng acceleration, rh/accelerate.c:34-146, on the flattened populations */
typedef struct { int N, Norder, Nperiod, Ndelay, count; double *prev; } ng_t;

static void ng_init(ng_t *g, int N, int Ndelay, int Norder, int Nperiod, const double *sol)
{
  int k;
  g->N = N; g->Norder = Norder; g->Nperiod = Nperiod;
  g->Ndelay = MAXV(Ndelay, Norder + 2);
  g->prev = calloc((size_t) (Norder + 2) * N, sizeof(double));
  for (k = 0; k < N; k++) g->prev[k] = sol[k];
  g->count = 1;
}

static void ng_accelerate(ng_t *g, double *sol)
{
  const int Norder = g->Norder, N = g->N;
  int i, j, k, ip, ipp, i0;
  double A[NEQ_MAX][NEQ_MAX], b[NEQ_MAX], *Delta, *weight;
  double *cur = g->prev + (size_t) (g->count % (Norder + 2)) * N;
  for (k = 0; k < N; k++) cur[k] = sol[k];
  g->count++;
  if (!(Norder > 0 && g->count >= g->Ndelay && !((g->count - g->Ndelay) % g->Nperiod))) return;
  Delta = malloc((size_t) (Norder + 1) * N * sizeof(double));
  weight = malloc((size_t) N * sizeof(double));
  for (i = 0; i <= Norder; i++) {
    ip = (g->count - 1 - i) % (Norder + 2);
    ipp = (g->count - 2 - i) % (Norder + 2);
    for (k = 0; k < N; k++)
      Delta[(size_t) i * N + k] = g->prev[(size_t) ip * N + k] - g->prev[(size_t) ipp * N + k];
  }
  for (k = 0; k < N; k++) weight[k] = 1.0 / fabs(sol[k]);
  for (i = 0; i < Norder; i++) { b[i] = 0.0; for (j = 0; j < Norder; j++) A[i][j] = 0.0; }
  for (j = 0; j < Norder; j++) {
    for (k = 0; k < N; k++)
      b[j] += weight[k] * Delta[k] * (Delta[k] - Delta[(size_t) (j + 1) * N + k]);
    for (i = 0; i < Norder; i++)
      for (k = 0; k < N; k++)
        A[i][j] += weight[k] * (Delta[(size_t) (j + 1) * N + k] - Delta[k]) *
                   (Delta[(size_t) (i + 1) * N + k] - Delta[k]);
  }
  solve_linear_improve(Norder, A, b);
  i0 = (g->count - 1) % (Norder + 2);
  for (i = 0; i < Norder; i++) {
    ip = (g->count - 2 - i) % (Norder + 2);
    for (k = 0; k < N; k++)
      sol[k] += b[i] * (g->prev[(size_t) ip * N + k] - g->prev[(size_t) i0 * N + k]);
  }
  for (k = 0; k < N; k++) g->prev[(size_t) i0 * N + k] = sol[k];
  free(weight);
  free(Delta);
}

/* MaxChange, rh/maxchange.c:32-50 */
static double ng_maxchange(const ng_t *g)
{
  double dmax = 0.0;
  const double *old, *cur;
  int k;
  if (g->count < 2) return dmax;
  old = g->prev + (size_t) ((g->count - 2) % (g->Norder + 2)) * g->N;
  cur = g->prev + (size_t) ((g->count - 1) % (g->Norder + 2)) * g->N;
  for (k = 0; k < g->N; k++)
    if (cur[k]) dmax = MAXV(dmax, fabs((cur[k] - old[k]) / cur[k]));
  return dmax;
}

/* Intermediate arrays of the first column solved, written in the capture
   record format of ref_harness.c when RHO_NLTE_DUMP names a file (tests
   compare them with the reference's). */
static FILE *nlte_dump_open(void)
{
  static int done = 0;
  const char *p = getenv("RHO_NLTE_DUMP");
  FILE *f = NULL;
#ifdef _OPENMP
#pragma omp critical(nlte_dump)
#endif
  {
    if (p && !done) { f = fopen(p, "wb"); done = 1; }
  }
  return f;
}

static void nlte_dump(FILE *f, const char *tag, int idx, const double *x, int n)
{
  char t[16];
  int hdr[5] = {0, idx, 0, 0, n};
  if (!f) return;
  memset(t, 0, sizeof(t));
  strncpy(t, tag, 15);
  fwrite(t, 1, 16, f);
  fwrite(hdr, sizeof(int), 5, f);
  fwrite(x, sizeof(double), n, f);
}

static int nlte_column(const rhg_model_t *m, const rhg_nlte_t *na, const column_t *c,
                       struct nlte_out *o)
{
  FILE *dump = nlte_dump_open();
  const int N = c->Ndep, nl = na->nlevel, nspect = m->nspect, nrays = na->nrays;
  const int nlev_all = m->atom_level_ptr[m->natom], l0 = m->atom_level_ptr[na->atom_index];
  const double vtherm = 2.0 * KBOLTZMANN / (AMU * na->weight);
  int k, i, l, t, ls, mu, to_obs, la, niter, maxset = 1, *phi_off;
  size_t nphi = 0;
  double *adamp, dpops = 0.0;
  nlte_ws w;
  ng_t ng;
  if (nl > NLEV_MAX) return -1;
  memset(&w, 0, sizeof(w));
  w.m = m; w.na = na; w.c = c; w.N = N; w.nl = nl; w.nrays = nrays;
  for (ls = 0; ls < nspect; ls++) maxset = MAXV(maxset, na->as_ptr[ls + 1] - na->as_ptr[ls]);
  phi_off = calloc((size_t) na->nline + 1, sizeof(int));
  for (l = 0; l < na->nline; l++) {
    phi_off[l] = (int) nphi;
    for (t = 0; t < na->ntrans; t++)
      if (na->tr_line[t] == l) nphi += (size_t) na->tr_nlam[t] * nrays * 2 * N;
  }
  w.phi_off = phi_off;
  w.n = calloc((size_t) nl * N, sizeof(double));
  w.nstar = calloc((size_t) nl * N, sizeof(double));
  w.ntotal = calloc(N, sizeof(double));
  w.vbroad = calloc(N, sizeof(double));
  w.C = calloc((size_t) nl * nl * N, sizeof(double));
  w.Gamma = calloc((size_t) nl * nl * N, sizeof(double));
  w.J = calloc((size_t) nspect * N, sizeof(double));
  w.Jdag = calloc(N, sizeof(double));
  w.phi = calloc(nphi + 1, sizeof(double));
  w.wphi = calloc((size_t) (na->nline + 1) * N, sizeof(double));
  w.bg_chi = calloc((size_t) nspect * N, sizeof(double));
  w.bg_eta = calloc((size_t) nspect * N, sizeof(double));
  w.bg_sca = calloc((size_t) nspect * N, sizeof(double));
  w.chi = calloc(N, sizeof(double)); w.S = calloc(N, sizeof(double));
  w.I = calloc(N, sizeof(double)); w.Psi = calloc(N, sizeof(double));
  w.eta_at = calloc(N, sizeof(double)); w.Ieff = calloc(N, sizeof(double));
  w.scr = calloc((size_t) 8 * N, sizeof(double));
  w.V = calloc((size_t) maxset * N, sizeof(double));
  w.gij = calloc((size_t) maxset * N, sizeof(double));
  w.wla = calloc((size_t) maxset * N, sizeof(double));
  w.chi_up = calloc((size_t) nl * N, sizeof(double));
  w.chi_down = calloc((size_t) nl * N, sizeof(double));
  w.Uji = calloc((size_t) nl * N, sizeof(double));
  adamp = calloc(N, sizeof(double));
  w.adamp_l = calloc((size_t) (na->nline + 1) * N, sizeof(double));

  /* LTE populations and Doppler widths (rh/readatom.c:213-220) */
  for (i = 0; i < nl; i++)
    for (k = 0; k < N; k++) w.nstar[(size_t) i * N + k] = c->nstar[(size_t) k * nlev_all + l0 + i];
  for (k = 0; k < N; k++) {
    w.ntotal[k] = m->atom_abund[na->atom_index] * c->nHtot[k];
    w.vbroad[k] = sqrt(vtherm * c->T[k] + SQ(c->vturb[k]));
  }

  /* line profiles and their normalisation (rh/profile.c:67-371 with
     Damping, rh/broad.c:273-316) */
  for (l = 0; l < na->nline; l++) {
    const double lambda0 = na->ln_lambda0[l], cDop = (NM_TO_M * lambda0) / (4.0 * PI);
    double *wphi = w.wphi + (size_t) l * N;
    for (t = 0; t < na->ntrans && na->tr_line[t] != l; t++);
    for (k = 0; k < N; k++) {
      double Qelast = 0.0, G;
      if (na->ln_vdw[l] == 2) {   /* rh/broad.c:125-136 */
        G = na->ln_vdw_cross[l] * pow(c->T[k], (1.0 - na->ln_vdw_alpha[l]) / 2.0) +
            na->ln_vdw_cross_he[l] * pow(c->T[k], 0.3);
        G *= c->nH0[k];
        Qelast += G;
      } else if (na->ln_vdw[l]) { /* rh/broad.c:87-100 */
        G = na->ln_vdw_cross[l] * pow(c->T[k], 0.3);
        G *= c->nH0[k];
        Qelast += G;
      }
      if (na->ln_stark[l] == 1) {
        Qelast += na->ln_stark_c[l] * c->ne[k];
      } else if (na->ln_stark[l] == 2) {
        const double vrel = pow(na->ln_stark_C[l] * c->T[k], 0.16666667) * na->ln_stark_Cm[l];
        Qelast += na->ln_stark_c[l] * vrel * c->ne[k];
      }
      adamp[k] = (na->ln_Grad[l] + Qelast) * cDop / w.vbroad[k];
      wphi[k] = 0.0;
    }
    nlte_dump(dump, "adamp", l, adamp, N);
    for (k = 0; k < N; k++) w.adamp_l[(size_t) l * N + k] = adamp[k];
    for (la = 0; la < na->tr_nlam[t]; la++) {
      const double lamla = m_lambda(m, na->tr_nblue[t] + la);
      const double wl = na->tr_wlam[na->tr_ptr[t] + la];
      for (mu = 0; mu < nrays; mu++) {
        const double wlamu = wl * 0.5 * na->wmu[mu];
        for (to_obs = 0; to_obs <= 1; to_obs++) {
          const double sign = to_obs ? 1.0 : -1.0;
          double *phi = w.phi + phi_off[l] + ((size_t) (la * nrays + mu) * 2 + to_obs) * N;
          for (k = 0; k < N; k++) {
            const double v = (lamla - lambda0 - 0.0) * CLIGHT / (w.vbroad[k] * lambda0);
            const double vlos = na->muz[mu] * c->vz[k] / w.vbroad[k];
            const double vk = v + sign * vlos;
            phi[k] = 0.0;
            phi[k] += rho_voigt_armstrong(adamp[k], vk) * 1.0 / (SQRTPI * w.vbroad[k]);
            wphi[k] += phi[k] * wlamu;
          }
        }
      }
    }
    for (k = 0; k < N; k++) wphi[k] = 1.0 / wphi[k];
  }

  /* collisional rates (rh/collision.c:450-760) */
  {
    const double C0 = ((E_RYDBERG / sqrt(M_ELECTRON)) * PI * SQ(RBOHR)) * sqrt(8.0 / (PI * KBOLTZMANN));
    int r;
    for (r = 0; r < na->ncoll; r++) {
      const int i1 = na->coll_i[r], j1 = na->coll_j[r], nT = na->coll_ptr[r + 1] - na->coll_ptr[r];
      const double *Tt = na->coll_T + na->coll_ptr[r], *ct = na->coll_c + na->coll_ptr[r],
                   *Mt = na->coll_M + na->coll_ptr[r];
      double *Cij = w.C + (size_t) (i1 * nl + j1) * N, *Cji = w.C + (size_t) (j1 * nl + i1) * N;
      const double *nsi = w.nstar + (size_t) i1 * N, *nsj = w.nstar + (size_t) j1 * N;
      for (k = 0; k < N; k++) {
        const double T = c->T[k];
        const double Cv = nT > 2 ? spline_eval_M(nT, Tt, ct, Mt, T) : linear1(nT, Tt, ct, T);
        double Cdown, Cup;
        switch (na->coll_type[r]) {
        case 0:
          Cdown = C0 * c->ne[k] * Cv / (na->g[j1] * sqrt(T));
          Cij[k] += Cdown; Cji[k] += Cdown * nsj[k] / nsi[k];
          break;
        case 1:
          Cdown = Cv * c->ne[k] * (na->g[i1] / na->g[j1]) * sqrt(T);
          Cij[k] += Cdown; Cji[k] += Cdown * nsj[k] / nsi[k];
          break;
        case 2:
          Cup = Cv * c->ne[k] * exp(-(na->E[j1] - na->E[i1]) / (KBOLTZMANN * T)) * sqrt(T);
          Cji[k] += Cup; Cij[k] += Cup * nsi[k] / nsj[k];
          break;
        case 3:
          Cdown = c->np[k] * Cv;
          Cij[k] += Cdown; Cji[k] += Cdown * nsj[k] / nsi[k];
          break;
        case 4:
          Cup = c->nH0[k] * Cv;
          Cji[k] += Cup; Cij[k] += Cup * nsi[k] / nsj[k];
          break;
        case 5: Cij[k] += c->nH0[k] * Cv; break;
        case 6: Cji[k] += c->np[k] * Cv; break;
        }
      }
    }
  }

  /* background at the wavelengths with an active set: the LTE continuum
     and the scattering opacity (rh/background.c:342-465), which NLTE adds
     to chi and takes J-weighted into the source function */
  for (ls = 0; ls < nspect; ls++) {
    double *bc = w.bg_chi + (size_t) ls * N, *be = w.bg_eta + (size_t) ls * N,
           *bs = w.bg_sca + (size_t) ls * N;
    if (ls == m->ref_index) continue;
    background_continuum(m, c, ls, bc, be);
    for (k = 0; k < N; k++) {
      double sca = c->ne[k] * na->sigma_thomson;
      sca += na->ray_H[ls] * c->nH0[k];
      if (na->he_atom >= 0)
        sca += na->ray_He[ls] * c->nstar[(size_t) k * nlev_all + m->atom_level_ptr[na->he_atom]];
      sca += na->ray_H2[ls] * c->nH2[k];
      /* scattering fudge factor, rh/background.c:453-462 */
      sca *= m->lamconst[(size_t) ls * RHG_LC_NCONST + RHG_LC_SCATT_FUDGE];
      bs[k] = sca;
      bc[k] += sca;
    }
  }

  for (ls = 0; ls < nspect; ls++) {
    nlte_dump(dump, "bgc_chi", ls, w.bg_chi + (size_t) ls * N, N);
    nlte_dump(dump, "bgc_eta", ls, w.bg_eta + (size_t) ls * N, N);
    nlte_dump(dump, "bgc_sca", ls, w.bg_sca + (size_t) ls * N, N);
  }
  for (i = 0; i < nl * nl; i++) nlte_dump(dump, "coll_C", i, w.C + (size_t) i * N, N);
  for (i = 0; i < nl; i++) nlte_dump(dump, "nlte_nstar", i, w.nstar + (size_t) i * N, N);
  for (l = 0; l < na->nline; l++) {
    nlte_dump(dump, "wphi", l, w.wphi + (size_t) l * N, N);
    nlte_dump(dump, "phi0", l, w.phi + phi_off[l], N);
  }

  w.dump = dump;
  /* initial solution (LTE populations, J = 0) and background scattering
     iterations (rh/initial_xdr.c:276-282, rh/initscatter.c:60-75) */
  for (i = 0; i < nl * N; i++) w.n[i] = w.nstar[i];
  if (na->nmax_iter > 0) {
    int it = 0;
    while (it < na->nmax_scatter) {
      double dJmax = 0.0;
      for (ls = 0; ls < nspect; ls++)
        if (ls != m->ref_index) {
          const double dJ = nlte_formal(&w, ls, 0);
          dJmax = MAXV(dJmax, dJ);
        }
      w.dump = NULL;
      if (dJmax < na->iter_limit) break;
      it++;
    }
  }

  for (ls = 0; ls < nspect; ls++) nlte_dump(dump, "J_scat", ls, w.J + (size_t) ls * N, N);

  /* main iterations (rh/iterate.c:48-143, rh/statequil.c:36-216) */
  ng_init(&ng, nl * N, na->ng_delay, na->ng_order, na->ng_period, w.n);
  for (niter = 1; niter <= na->nmax_iter; niter++) {
    for (i = 0; i < nl * nl * N; i++) w.Gamma[i] = w.C[i];
    for (ls = 0; ls < nspect; ls++)
      if (ls != m->ref_index) nlte_formal(&w, ls, 1);
    if (niter == 1)
      for (i = 0; i < nl * nl; i++) nlte_dump(dump, "gamma1", i, w.Gamma + (size_t) i * N, N);
    for (k = 0; k < N; k++) {
      double A[NEQ_MAX][NEQ_MAX], nk[NEQ_MAX], nmax = 0.0;
      int j, ie = 0;
      for (i = 0; i < nl; i++) {
        nk[i] = w.n[(size_t) i * N + k];
        for (j = 0; j < nl; j++) A[i][j] = w.Gamma[(size_t) (i * nl + j) * N + k];
      }
      if (na->isum == -1) {
        for (i = 0; i < nl; i++) if (nk[i] > nmax) { nmax = nk[i]; ie = i; }
      } else
        ie = na->isum;
      for (i = 0; i < nl; i++) {
        double GamDiag = 0.0;
        A[i][i] = 0.0;
        nk[i] = 0.0;
        for (j = 0; j < nl; j++) GamDiag += A[j][i];
        A[i][i] = -GamDiag;
      }
      nk[ie] = w.ntotal[k];
      for (j = 0; j < nl; j++) A[ie][j] = 1.0;
      solve_linear_improve(nl, A, nk);
      for (i = 0; i < nl; i++) w.n[(size_t) i * N + k] = nk[i];
    }
    ng_accelerate(&ng, w.n);
    dpops = ng_maxchange(&ng);
    if (dpops < na->iter_limit) break;
  }
  if (niter > na->nmax_iter) niter = na->nmax_iter;

  /* J with the converged populations: up to N_MAX_SCATTER passes without
     the operator, stopping at max dJ <= ITER_LIMIT
     (rh/rhf1d/pyrh_compute1dray.c:332-337) */
  if (o->stokes) {
    int it = 0;
    while (it < na->nmax_scatter) {
      double dJmax = 0.0;
      for (ls = 0; ls < nspect; ls++)
        if (ls != m->ref_index) {
          const double dJ = nlte_formal(&w, ls, 0);
          dJmax = MAXV(dJmax, dJ);
        }
      if (dJmax <= na->iter_limit) break;
      it++;
    }
  }

  /* the emergent spectrum at the column's mu with the converged
     populations and the last J (pyrh's _solveray, rh/rhf1d/pyrh_solveray.c:
     82-107: one ray of weight 1, profiles for it, background scattering
     from the stored J, no update of J; rh/rhf1d/formal.c:44-346) */
  if (o->stokes) {
    const double muf = o->mu;
    int io = 0;
    for (ls = 0; ls < nspect; ls++) {
      const double lam = m_lambda(m, ls);
      const double *chic = w.bg_chi + (size_t) ls * N, *etac = w.bg_eta + (size_t) ls * N,
                   *scac = w.bg_sca + (size_t) ls * N, *Jl = w.J + (size_t) ls * N;
      const int p0 = na->as_ptr[ls], p1 = na->as_ptr[ls + 1];
      double B2[2], Iout;
      int e;
      if (ls == m->ref_index) continue;
      for (k = 0; k < N; k++) { w.chi[k] = 0.0; w.eta_at[k] = 0.0; }
      for (e = p0; e < p1; e++) {
        const int t = na->as_tr[e], it = na->tr_i[t], jt = na->tr_j[t];
        const int la2 = ls - na->tr_nblue[t];
        const double *ni = w.n + (size_t) it * N, *nj = w.n + (size_t) jt * N;
        for (k = 0; k < N; k++) {
          double V, gij, tw;
          if (na->tr_type[t] == 0) {
            const int l = na->tr_line[t];
            const double lambda0 = na->ln_lambda0[l];
            const double v = (lam - lambda0 - 0.0) * CLIGHT / (w.vbroad[k] * lambda0);
            const double vlos = muf * c->vz[k] / w.vbroad[k];
            const double ph = rho_voigt_armstrong(w.adamp_l[(size_t) l * N + k], v + 1.0 * vlos) *
                              1.0 / (SQRTPI * w.vbroad[k]);
            gij = na->ln_Bji[l] / na->ln_Bij[l];
            V = (HPLANCK * CLIGHT / (4.0 * PI) * na->ln_Bij[l] * 1.0) * ph;
            tw = na->ln_Aji[l] / na->ln_Bji[l];
          } else {
            tw = 2.0 * HPLANCK * CLIGHT / CUBE(NM_TO_M) / CUBE(lam);
            V = na->tr_alpha[na->tr_ptr[t] + la2];
            gij = w.nstar[(size_t) it * N + k] / w.nstar[(size_t) jt * N + k] *
                  exp(-HPLANCK * CLIGHT / (KBOLTZMANN * NM_TO_M) / (lam * c->T[k]));
          }
          w.chi[k] += V * (ni[k] - gij * nj[k]);
          w.eta_at[k] += tw * gij * V * nj[k];
        }
      }
      B2[0] = rho_planck(c->T[N - 2], lam);
      B2[1] = rho_planck(c->T[N - 1], lam);
      if (na->as_bezier[ls]) {
        for (k = 0; k < N; k++) {
          const double chi_as = w.chi[k];
          w.chi[k] = chi_as + chic[k];
          w.S[k] = w.eta_at[k] + etac[k];
          w.S[k] += scac[k] * Jl[k];
        }
        for (k = 0; k < N; k++) w.S[k] /= w.chi[k];
        bezier_ray(N, c->height, 1.0 / muf, w.chi, w.S, B2, 1, w.I, w.Psi);
        nlte_dump(dump, "fin_chi", ls, w.chi, N);
        nlte_dump(dump, "fin_S", ls, w.S, N);
        nlte_dump(dump, "fin_J", ls, Jl, N);
        Iout = w.I[0];
      } else {
        for (k = 0; k < N; k++) {
          const double chi_as = w.chi[k];
          w.chi[k] = chi_as + chic[k];
          w.S[k] = (w.eta_at[k] + etac[k] + scac[k] * Jl[k]) / w.chi[k];
        }
        Iout = rho_formal_feautrier(N, c->height, muf, w.chi, w.S, B2);
      }
      o->stokes[0 * (nspect - 1) + io] = Iout;
      o->stokes[1 * (nspect - 1) + io] = 0.0;
      o->stokes[2 * (nspect - 1) + io] = 0.0;
      o->stokes[3 * (nspect - 1) + io] = 0.0;
      io++;
    }
  }
  for (i = 0; i < nl; i++)
    for (k = 0; k < o->ndep_max; k++)
      o->pops[(size_t) i * o->ndep_max + k] = k < N ? w.n[(size_t) i * N + k] : 0.0;
  *o->niter = niter;
  *o->dpops = dpops;

  if (dump) fclose(dump);
  free(ng.prev);
  free(adamp); free(phi_off); free(w.adamp_l);
  free(w.n); free(w.nstar); free(w.ntotal); free(w.vbroad); free(w.C); free(w.Gamma);
  free(w.J); free(w.Jdag); free(w.phi); free(w.wphi); free(w.bg_chi); free(w.bg_eta);
  free(w.bg_sca); free(w.chi); free(w.S); free(w.I); free(w.Psi); free(w.eta_at);
  free(w.Ieff); free(w.scr); free(w.V); free(w.gij); free(w.wla); free(w.chi_up);
  free(w.chi_down); free(w.Uji);
  return 0;
}

/* NLTE populations of the active atom per column: pops_out
   [ncol][nlevel][ndep_max], iterations and final max relative change */
int rho_nlte_solve(const rhg_model_t *m, const rhg_nlte_t *na, const rhg_columns_t *cols,
                   int nthreads, double *pops_out, int *niter_out, double *dpops_out,
                   double *stokes_out)
{
  int ic, err = 0;
  if (m->nnuclei + m->nmol > NEQ_MAX) return RHG_EUNSUPPORTED;
  if (na->nlevel > NLEV_MAX || na->ng_order + 2 > NEQ_MAX) return RHG_EUNSUPPORTED;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(|:err)
#endif
  for (ic = 0; ic < cols->ncol; ic++) {
    struct nlte_out o;
    o.pops = pops_out + (size_t) ic * na->nlevel * cols->ndep_max;
    o.niter = niter_out + ic;
    o.dpops = dpops_out + ic;
    o.ndep_max = cols->ndep_max;
    o.stokes = stokes_out ? stokes_out + (size_t) ic * 4 * (m->nspect - 1) : NULL;
    o.mu = cols->mu;
    if (synth_column_ex(m, cols, ic, 0, NULL, NULL, NULL, -1, NULL, NULL, NULL, 0, NULL, NULL,
                        na, &o))
      err |= 1;
  }
  (void) nthreads;
  return err ? RHG_ENOMEM : RHG_OK;
}

/* layout check for the host bindings */
int rho_model_sizeof(void) { return (int) sizeof(rhg_model_t); }
int rho_columns_sizeof(void) { return (int) sizeof(rhg_columns_t); }

/* ------------------------------------------------------------------ */
/* hydrostatic equilibrium (pyrh's hse())                              */

#define HSE_MAX_ELECTRON_ERROR 1.0E-2     /* rh/rhf1d/pyrh_background.h:1-2 */
#define HSE_N_MAX_ELECTRON_ITER 10

/* getfjk, rh/solvene.c:145-197 (LTE branch: Kurucz partition functions,
   getKuruczpf rh/solvene.c:202-211): ionisation fractions f_j and their
   derivatives with ne for element e of the table */
static void hse_getfjk(const rhg_model_t *m, const rhg_hse_t *h, int e, double T,
                       double ne, double *fjk, double *dfjk)
{
  const double C1 = (HPLANCK / (2.0 * PI * M_ELECTRON)) * (HPLANCK / KBOLTZMANN);
  const int ns = h->el_nstage[e], npf = m->npf;
  const double *lp = h->el_logpf + h->el_pf_ptr[e], *ip = h->el_ionpot + h->el_pf_ptr[e];
  double CT_ne = 2.0 * pow(C1 / T, -1.5) / ne, sum1 = 1.0, sum2 = 0.0, Uk, Ukp1;
  int j;
  fjk[0] = 1.0;
  dfjk[0] = 0.0;
  Uk = linear1(npf, m->Tpf, lp, T);
  for (j = 1; j < ns; j++) {
    Ukp1 = linear1(npf, m->Tpf, lp + j * npf, T);
    fjk[j] = fjk[j - 1] * CT_ne * exp(Ukp1 - Uk - ip[j - 1] / (KBOLTZMANN * T));
    dfjk[j] = -j * fjk[j] / ne;
    sum1 += fjk[j];
    sum2 += dfjk[j];
    Uk = Ukp1;
  }
  for (j = 0; j < ns; j++) {
    fjk[j] /= sum1;
    dfjk[j] = (dfjk[j] - fjk[j] * sum2) / sum1;
  }
}

/* get_ne(fromscratch = TRUE) with hydrogen in LTE,
   rh/rhf1d/pyrh_background.c:459-549: starting value from hydrogen
   ionisation with U = Uk (0 there; Solve_ne, rh/solvene.c:78-88, takes
   the Kurucz partition function of neutral hydrogen), Newton steps over
   all elements and H- (the same in both). */
static _Thread_local long long g_newton_iter = 0;

static double hse_get_ne(const rhg_model_t *m, const rhg_hse_t *h, double T, double nHtot,
                         double Uk)
{
  const double C1 = (HPLANCK / (2.0 * PI * M_ELECTRON)) * (HPLANCK / KBOLTZMANN);
  double fjk[NSTAGE_MAX], dfjk[NSTAGE_MAX], error, sum, akj, PhiH, PhiHmin, ne, ne_old, dne;
  int niter = 0, n, j;
  PhiH = 0.5 * pow(C1 / T, 1.5) * exp(Uk + h->el_ionpot[h->el_pf_ptr[0]] / (KBOLTZMANN * T));
  ne_old = (sqrt(1.0 + 4.0 * nHtot * PhiH) - 1.0) / (2.0 * PhiH);
  ne = ne_old;
  while (niter < HSE_N_MAX_ELECTRON_ITER) {
    error = ne_old / nHtot;
    sum = 0.0;
    for (n = 0; n < h->nel; n++) {
      hse_getfjk(m, h, n, T, ne_old, fjk, dfjk);
      if (n == 0) {
        PhiHmin = 0.25 * pow(C1 / T, 1.5) * exp(0.754 * EV / (KBOLTZMANN * T));
        error += ne_old * fjk[0] * PhiHmin;
        sum -= (fjk[0] + ne_old * dfjk[0]) * PhiHmin;
      }
      for (j = 1; j < h->el_nstage[n]; j++) {
        akj = h->el_abund[n] * j;
        error -= akj * fjk[j];
        sum += akj * dfjk[j];
      }
    }
    ne = ne_old - nHtot * error / (1.0 - nHtot * sum);
    dne = fabs((ne - ne_old) / ne_old);
    ne_old = ne;
    g_newton_iter++;
    if (dne <= HSE_MAX_ELECTRON_ERROR) break;
    niter++;
  }
  return ne;
}

/* Continuum opacity at wavelength ls in one layer as pyrh_Background sums
   it, rh/rhf1d/pyrh_background.c:256-400: H- bound-free and free-free
   (times the H- fudge factor), OH and CH bound-free, H bound-free and
   free-free, H2+ and H2- free-free; plus the scattering (Thomson, Rayleigh
   on H, He and H2, times the scattering fudge factor).  No metal
   bound-free (commented out there).  ns: all level populations. */
static double hse_opacity(const rhg_model_t *m, const rhg_hse_t *h, int ls, double T, double ne,
                          const double *ns, double nHmin, const double *nmol)
{
  const double *lc = m->lamconst + (size_t) ls * RHG_LC_NCONST;
  const int lH = m->atom_level_ptr[0], nlH = m->atom_level_ptr[1] - lH;
  const double nH0 = ns[lH], np = ns[lH + nlH - 1];
  const double nH2 = m->mol_h2 >= 0 ? nmol[m->mol_h2] : 0.0;
  const double nOH = m->mol_oh >= 0 ? nmol[m->mol_oh] : 0.0;
  const double nCH = m->mol_ch >= 0 ? nmol[m->mol_ch] : 0.0;
  double C0 = SQ(Q_ELECTRON) / (4.0 * PI * EPSILON_0) / sqrt(M_ELECTRON);
  double sigma_ff = 4.0 / 3.0 * sqrt(2.0 * PI / (3.0 * KBOLTZMANN)) * CUBE(C0) /
                    (HPLANCK * CLIGHT);
  double total = 0.0, sca, chi, stim;
  int e, cp0 = m->lam_cont_ptr[ls], cp1 = m->lam_cont_ptr[ls + 1];
  /* Thomson, rh/thomson.c:30-48 */
  sca = ne * h->sigma_thomson;
  if (lc[RHG_LC_HMIN_BF_ON] != 0.0) {
    stim = exp(-lc[RHG_LC_HC_KLA] / T);
    total += nHmin * (1.0 - stim) * lc[RHG_LC_HMIN_BF_ALPHA];
  }
  if (lc[RHG_LC_HMIN_FF_LIDX] >= 0.0) {
    double ti = clamp_index(NTHETA, m->hminff_theta, THETA0 / T);
    double pe = ne * KBOLTZMANN * T;
    double kappa = bilinear(NTHETA, NFF, m->hminff_kappa, ti, lc[RHG_LC_HMIN_FF_LIDX]);
    chi = (nH0 * 1.0E-29) * pe * kappa;
  } else {
    double Ck = (KBOLTZMANN * THETA0 * 1.0E-32), theta_n = 1.0, sqrt_theta = sqrt(THETA0 / T);
    int n;
    chi = 0.0;
    for (n = 1; n < 6; n++) {
      theta_n *= sqrt_theta;
      chi += theta_n * lc[RHG_LC_HFF_LONG0 + n - 1];
    }
    chi *= nH0 * (ne * Ck);
  }
  total += chi;
  total *= h->hmin_fudge;
  if (lc[RHG_LC_OH_EIDX] >= 0.0) {
    if (T < m->oh_T[0] || T > m->oh_T[NTOH - 1]) chi = 0.0;
    else {
      int i2 = locate_asc(NTOH, m->oh_T, T);
      double ti = (i2 == NTOH - 1) ? (double) i2 :
        (double) i2 + (T - m->oh_T[i2]) / (m->oh_T[i2 + 1] - m->oh_T[i2]);
      double kappa = exp(LG10 * bilinear(NTOH, NEOH, m->oh_cross, ti,
                                         lc[RHG_LC_OH_EIDX])) * SQ(CM_TO_M);
      stim = exp(-lc[RHG_LC_HC_KLA] / T);
      chi = nOH * (1.0 - stim) * kappa;
    }
    total += chi;
  }
  if (lc[RHG_LC_CH_EIDX] >= 0.0) {
    if (T < m->ch_T[0] || T > m->ch_T[NTCH - 1]) chi = 0.0;
    else {
      int i2 = locate_asc(NTCH, m->ch_T, T);
      double ti = (i2 == NTCH - 1) ? (double) i2 :
        (double) i2 + (T - m->ch_T[i2]) / (m->ch_T[i2 + 1] - m->ch_T[i2]);
      double kappa = exp(LG10 * bilinear(NTCH, NECH, m->ch_cross, ti,
                                         lc[RHG_LC_CH_EIDX])) * SQ(CM_TO_M);
      stim = exp(-lc[RHG_LC_HC_KLA] / T);
      chi = nCH * (1.0 - stim) * kappa;
    }
    total += chi;
  }
  /* H bound-free, rh/hydrogen.c:150-208 (active-layer branch) */
  {
    int any = 0;
    chi = 0.0;
    for (e = cp0; e < cp1; e++) {
      int ci = m->lam_cont_idx[e];
      if (m->cont_atom[ci] != 0) continue;
      any = 1;
      {
        double expla = exp(-lc[RHG_LC_BF_HC_KLA] / T);
        chi += m->lam_cont_alpha[e] * (1.0 - expla) * ns[lH + m->cont_i[ci]];
      }
    }
    if (any) total += chi;
  }
  {
    double y = (lc[RHG_LC_GFF_CY] * T) / (HPLANCK * CLIGHT), x3 = lc[RHG_LC_GFF_X3], g_ff;
    double gIII = 1.0 + 0.1728 * x3 * (1.0 + y) -
                  0.0496 * SQ(x3) * (1.0 + (1.0 + y) * 0.33333333 * y);
    g_ff = (gIII > 1.0) ? gIII : 1.0;
    stim = 1.0 - exp(-lc[RHG_LC_HC_KLA] / T);
    total += sigma_ff / sqrt(T) * lc[RHG_LC_HFF_NU3] * ne * np * stim * g_ff;
  }
  /* Rayleigh on H and He, rh/rayleigh.c:35-100 */
  if (h->ray_H > 0.0) sca += h->ray_H * nH0;
  if (h->he_atom >= 0 && h->ray_He > 0.0) sca += h->ray_He * ns[m->atom_level_ptr[h->he_atom]];
  if (lc[RHG_LC_H2P_LIDX] >= 0.0) {
    double ti = clamp_index(NTEMP_H2P, m->h2pff_temp, T);
    double kappa = bilinear(NTEMP_H2P, NFF_H2P, m->h2pff_kappa, ti, lc[RHG_LC_H2P_LIDX]);
    total += (nH0 * 1.0E-29) * (np * 1.0E-20) * kappa;
  }
  /* Rayleigh on H2, rh/hydrogen.c:938-995 */
  if (h->ray_H2 > 0.0) sca += h->ray_H2 * nH2;
  if (lc[RHG_LC_H2M_LIDX] >= 0.0) {
    if (nH2 > 0.0) {
      double ti = clamp_index(NTHETA_H2, m->h2mff_theta, THETA0 / T);
      double pe = ne * KBOLTZMANN * T;
      double kappa = bilinear(NTHETA_H2, NFF_H2, m->h2mff_kappa, ti, lc[RHG_LC_H2M_LIDX]);
      chi = (nH2 * 1.0E-29) * pe * kappa;
    } else chi = 0.0;
    total += chi;
  }
  sca *= h->scatt_fudge;
  return total + sca;
}

/* One pass of the layer loop body, rh/rhf1d/pyrh_hse.c:226-238 and
   :295-307: ne from get_ne, LTE populations of every model atom
   (SetLTEQuantities, rh/ltepops.c:222-249), chemical equilibrium and the
   opacity.  Scratch: ns [nlevel], ntot [natom], nmol [nmol]. */
static void hse_layer(const rhg_model_t *m, const rhg_hse_t *h, double T, double nHtot,
                      double *ns, double *ntot, double *nmol, double *ne_out, double *opac_out)
{
  double ne = hse_get_ne(m, h, T, nHtot, 0.0), nHmin;
  int ia;
  for (ia = 0; ia < m->natom; ia++) {
    ntot[ia] = m->atom_abund[ia] * nHtot;
    ltepops(m, ia, T, ne, ntot[ia], ns + m->atom_level_ptr[ia]);
  }
  nHmin = chemeq(m, T, ne, nHtot, ns, ntot, nmol);
  *ne_out = ne;
  *opac_out = hse_opacity(m, h, h->lam_index, T, ne, ns, nHmin, nmol);
}

/* pyrh's hse() for one column, rh/rhf1d/pyrh_hse.c:210-351 */
static int hse_column(const rhg_model_t *m, const rhg_hse_t *h, int N, int atm_scale,
                      const double *scale, const double *T, double pg_top, double *ne,
                      double *nHtot, double *rho, double *pg, int *niter)
{
  const double g = m->gravity, LOG10 = log(10.0);
  double *ns = malloc(sizeof(double) * (m->atom_level_ptr[m->natom] + m->natom + m->nmol + 2 * N + 1));
  double *ntot, *nmol, *opac, *z, nHtot_old, eta, deltaP = 0.0;
  int k, iter;
  if (!ns) return 1;
  ntot = ns + m->atom_level_ptr[m->natom];
  nmol = ntot + m->natom;
  opac = nmol + m->nmol;
  z = opac + N;
  for (k = 0; k < N; k++) z[k] = (atm_scale == 0) ? POW10(scale[k]) : scale[k] * KM_TO_M;
  pg[0] = pg_top;
  ne[0] = 0.0;
  nHtot[0] = (pg[0] / KBOLTZMANN / T[0] - ne[0]) / m->totalAbund;
  iter = 0;
  while (iter < h->nmax_iter) {
    rho[0] = (AMU * m->wght_per_H) * nHtot[0];
    hse_layer(m, h, T[0], nHtot[0], ns, ntot, nmol, ne, opac);
    nHtot_old = nHtot[0];
    nHtot[0] = (pg[0] / KBOLTZMANN / T[0] - ne[0]) / m->totalAbund;
    iter++;
    eta = fabs((nHtot[0] - nHtot_old) / nHtot[0]);
    if (eta <= h->iter_limit) break;
  }
  if (niter) niter[0] = iter;
  for (k = 1; k < N; k++) {
    iter = 0;
    if (atm_scale == 2) deltaP = g * rho[k - 1] * (z[k - 1] - z[k]);
    else deltaP = g * rho[k - 1] / opac[k - 1] * (z[k] - z[k - 1]);
    pg[k] = pg[k - 1] + deltaP;
    nHtot[k] = (pg[k] / KBOLTZMANN / T[k] - ne[k - 1]) / m->totalAbund;
    while (iter < h->nmax_iter) {
      rho[k] = (AMU * m->wght_per_H) * nHtot[k];
      hse_layer(m, h, T[k], nHtot[k], ns, ntot, nmol, ne + k, opac + k);
      if (atm_scale == 0) {
        double dlogtau = log10(z[k]) - log10(z[k - 1]);
        double beta1 = rho[k - 1] / opac[k - 1] * z[k - 1];
        double beta2 = rho[k] / opac[k] * z[k];
        pg[k] = pg[k - 1] + LOG10 * g * (beta2 + beta1) / 2 * dlogtau;
      } else {
        deltaP = g * (z[k - 1] - z[k]) * sqrt(rho[k] * rho[k - 1]);
        pg[k] = pg[k - 1] + deltaP;
      }
      nHtot_old = nHtot[k];
      nHtot[k] = (pg[k] / KBOLTZMANN / T[k] - ne[k]) / m->totalAbund;
      iter++;
      eta = fabs((nHtot[k] - nHtot_old) / nHtot[k]);
      if (eta <= h->iter_limit) break;
    }
    if (niter) niter[k] = iter;
  }
  free(ns);
  return 0;
}

int rho_hse(const rhg_model_t *m, const rhg_hse_t *h, int ncol, int ndep_max,
            const int *ndep, int atm_scale, const double *scale, const double *temp,
            const double *pg_top, int nthreads, double *ne, double *nHtot, double *rho,
            double *pg, int *niter)
{
  return rho_hse_census(m, h, ncol, ndep_max, ndep, atm_scale, scale, temp, pg_top, nthreads,
                        ne, nHtot, rho, pg, niter, NULL);
}

int rho_hse_census(const rhg_model_t *m, const rhg_hse_t *h, int ncol, int ndep_max,
                   const int *ndep, int atm_scale, const double *scale, const double *temp,
                   const double *pg_top, int nthreads, double *ne, double *nHtot, double *rho,
                   double *pg, int *niter, long long *census)
{
  int ic, err = 0;
  long long newton = 0, chem = 0;
  if (m->nnuclei + m->nmol > NEQ_MAX) return RHG_EUNSUPPORTED;
  if (atm_scale != 0 && atm_scale != 2) return RHG_EUNSUPPORTED;
  for (ic = 0; ic < h->nel; ic++) if (h->el_nstage[ic] > NSTAGE_MAX) return RHG_EUNSUPPORTED;
  memset(ne, 0, sizeof(double) * ncol * ndep_max);
  memset(nHtot, 0, sizeof(double) * ncol * ndep_max);
  memset(rho, 0, sizeof(double) * ncol * ndep_max);
  memset(pg, 0, sizeof(double) * ncol * ndep_max);
  if (niter) memset(niter, 0, sizeof(int) * ncol * ndep_max);
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(|:err) \
    reduction(+:newton, chem)
#endif
  for (ic = 0; ic < ncol; ic++) {
    size_t o = (size_t) ic * ndep_max;
    long long n0 = g_newton_iter, c0 = g_chem_iter;
    if (hse_column(m, h, ndep[ic], atm_scale, scale + o, temp + o, pg_top[ic], ne + o,
                   nHtot + o, rho + o, pg + o, niter ? niter + o : NULL))
      err |= 1;
    newton += g_newton_iter - n0;
    chem += g_chem_iter - c0;
  }
  (void) nthreads;
  if (census) {
    census[0] = newton;
    census[1] = chem;
  }
  return err ? RHG_ENOMEM : RHG_OK;
}

/* pyrh's get_ne_from_nH (rh/rhf1d/pyrh_hse.c:555-678): n_e per depth from
   T and n_Htot [cm^-3] by Solve_ne (rh/solvene.c:53-135) from scratch with
   hydrogen in LTE; returned in cm^-3 as the entry point converts back. */
int rho_ne_from_nh(const rhg_model_t *m, const rhg_hse_t *h, int ncol, int ndep_max,
                   const int *ndep, const double *temp, const double *nH, int nthreads,
                   double *ne_out)
{
  int ic;
  for (ic = 0; ic < h->nel; ic++) if (h->el_nstage[ic] > NSTAGE_MAX) return RHG_EUNSUPPORTED;
  memset(ne_out, 0, sizeof(double) * ncol * ndep_max);
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (ic = 0; ic < ncol; ic++) {
    size_t o = (size_t) ic * ndep_max;
    int k;
    for (k = 0; k < ndep[ic]; k++) {
      double T = temp[o + k], nHtot = nH[o + k] / CUBE(CM_TO_M);
      double Uk = linear1(m->npf, m->Tpf, h->el_logpf + h->el_pf_ptr[0], T);
      ne_out[o + k] = hse_get_ne(m, h, T, nHtot, Uk) * CUBE(CM_TO_M);
    }
  }
  (void) nthreads;
  return RHG_OK;
}

/* pyrh's get_scales (rh/rhf1d/pyrh_hse.c:402-553): the background
   opacity at the model's reference wavelength and convertScales
   (rh/rhf1d/multiatmos.c:100-177) for each column; tau, height [m] and
   cmass [kg m^-2] are [ncol][ndep_max], zero past ndep. */
int rho_scales(const rhg_model_t *m, const rhg_columns_t *cols, int nthreads, double *tau_out,
               double *height_out, double *cmass_out)
{
  int ic, err = 0;
  const int Nmax = cols->ndep_max;
  if (m->nnuclei + m->nmol > NEQ_MAX) return RHG_EUNSUPPORTED;
  memset(tau_out, 0, sizeof(double) * cols->ncol * Nmax);
  memset(height_out, 0, sizeof(double) * cols->ncol * Nmax);
  memset(cmass_out, 0, sizeof(double) * cols->ncol * Nmax);
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(|:err)
#endif
  for (ic = 0; ic < cols->ncol; ic++) {
    g_tau_out = tau_out + (size_t) ic * Nmax;
    g_cmass_out = cmass_out + (size_t) ic * Nmax;
    if (m->nspect != 1) { err |= 1; continue; }
    if (synth_column_ex(m, cols, ic, 0, NULL, height_out + (size_t) ic * Nmax,
                        NULL, -1, NULL, NULL, NULL, 0, NULL, NULL, NULL, NULL))
      err |= 1;
    g_tau_out = g_cmass_out = NULL;
  }
  (void) nthreads;
  return err ? RHG_ENOMEM : RHG_OK;
}
