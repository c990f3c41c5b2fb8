// The polarized cubic DELO-Bezier recursion of one emergent ray
// (rh/rhf1d/bezier_1D.c:52-300, to_obs: depth N-1 up to 0) split into its
// prologue, one depth step and the last (linear) interval, on an explicit
// state, for any number type T (double, float, or a dual number of either).
// The synthesis kernels run it on plain numbers; the response-function
// kernels run the same code on dual numbers to differentiate each step.
#pragma once

#include "rhg_kernels.hpp"

namespace rhg {

// Opacity, source functions and reduced propagation-matrix parameters of
// the emergent ray at one depth (rh/background.c:519-546 sums, formal.c:178-
// 208 source functions, stokesopac.c:28-88 K' with magneto-optical terms
// off: K' = [[0, q^T], [q, 0]]).
template <typename T>
struct Point {
  T chi;     // chi_I
  T S[4];    // eta_n / chi_I
  T q[3];    // chi_Q,U,V / chi_I
  T r[3];    // magneto-optical chip_Q,U,V / chi_I (MO only)
};

// Point from the continuum (cchi, ceta) and the line opacity at one depth;
// F(field) reads the column state there; the Humlicek functions in TV; MO:
// with the magneto-optical terms.
template <typename T, typename TV = T, bool MO = false, typename Field, typename Tab>
__device__ __forceinline__ void form_point_t(const Tab &tab, int ncont, int has_line,
                                             const Field &F, T mu, int moving, T cchi, T ceta,
                                             Point<T> &p) {
  T chi4[4], eta4[4], chip[3] = {T(0), T(0), T(0)};
  chi4[0] = cchi; eta4[0] = ceta;
  chi4[1] = chi4[2] = chi4[3] = T(0);
  eta4[1] = eta4[2] = eta4[3] = T(0);
  if (has_line) {
    T lI, lE, lq[3], le[3];
    line_point_t<T, TV, MO>(tab, ncont, F, mu, lI, lq, lE, le, moving, chip);
    chi4[0] += lI; eta4[0] += lE;
#pragma unroll
    for (int n = 0; n < 3; n++) { chi4[n + 1] += lq[n]; eta4[n + 1] += le[n]; }
  }
  p.chi = chi4[0];
  const T ichi = rrcp(chi4[0]);
#pragma unroll
  for (int n = 0; n < 4; n++) p.S[n] = eta4[n] * ichi;
#pragma unroll
  for (int n = 0; n < 3; n++) p.q[n] = chi4[n + 1] * ichi;
#pragma unroll
  for (int n = 0; n < 3; n++) p.r[n] = MO ? chip[n] * ichi : T(0);
}

// the same over wavelength ls's line list in the model
template <typename T, typename TV = T, bool MO = false, typename Field>
__device__ __forceinline__ void form_point_f(const DevModel &M, int ls, int has_line,
                                             const Field &F, T mu, int moving, T cchi, T ceta,
                                             Point<T> &p) {
  const int p0 = ldc(M.lam_line_ptr)[ls];
  form_point_t<T, TV, MO>(LineTabG(M, p0, ldc(M.lam_line_ptr)[ls + 1] - p0), M.ncont, has_line, F, mu,
                          moving, cchi, ceta, p);
}

// FP64 point over the column state at st (stride fs) with the Humlicek
// functions in TV and the continuum stored as R, from the line table tab
template <typename TV, typename R, bool MO = false, typename Tab>
__device__ __forceinline__ void form_point_vt(const Tab &tab, int ncont, int has_line,
                                              const double *st, int64_t fs, double mu, int moving,
                                              R cchi, R ceta, Point<double> &p) {
  auto F = [st, fs](int f) { return st[f * fs]; };
  form_point_t<double, TV, MO>(tab, ncont, has_line, F, mu, moving, (double) cchi, (double) ceta,
                               p);
}

// the same over wavelength ls's line list in the model
template <typename TV, typename R, bool MO = false>
__device__ __forceinline__ void form_point_v(const DevModel &M, int ls, int has_line,
                                             const double *st, int64_t fs, double mu, int moving,
                                             R cchi, R ceta, Point<double> &p) {
  const int p0 = ldc(M.lam_line_ptr)[ls];
  form_point_vt<TV, R, MO>(LineTabG(M, p0, ldc(M.lam_line_ptr)[ls + 1] - p0), M.ncont, has_line, st,
                           fs, mu, moving, cchi, ceta, p);
}

// y = K(q) x for K(q) = [[0, q^T], [q, 0]]
template <typename T>
__device__ __forceinline__ void kmul(const T *q, const T *x, T *y) {
  y[0] = q[0] * x[1] + q[1] * x[2] + q[2] * x[3];
  y[1] = q[0] * x[0];
  y[2] = q[1] * x[0];
  y[3] = q[2] * x[0];
}

// Solve (I + c K(q) + g (K(q)^2 + K(dq))) x = V.  The matrix is
// [[a, b^T], [b, I + g q q^T]] with a = 1 + g |q|^2 and b = c q + g dq; the
// lower block is inverted by Sherman-Morrison and the first component
// follows from its Schur complement.
template <typename T>
__device__ __forceinline__ void solve_md(T c, T g, const T *q, const T *dq, const T *V, T *x) {
  const T qq = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
  const T a = T(1) + g * qq;
  T b[3];
#pragma unroll
  for (int i = 0; i < 3; i++) b[i] = c * q[i] + g * dq[i];
  const T ga = rdiv(g, a);
  const T qb = q[0] * b[0] + q[1] * b[1] + q[2] * b[2];
  const T bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
  const T qV = q[0] * V[1] + q[1] * V[2] + q[2] * V[3];
  const T bV = b[0] * V[1] + b[1] * V[2] + b[2] * V[3];
  const T x0 = rdiv(V[0] - bV + ga * qV * qb, a - bb + ga * qb * qb);
  T y[3];
#pragma unroll
  for (int i = 0; i < 3; i++) y[i] = V[i + 1] - b[i] * x0;
  const T qy = ga * (q[0] * y[0] + q[1] * y[1] + q[2] * y[2]);
  x[0] = x0;
#pragma unroll
  for (int i = 0; i < 3; i++) x[i + 1] = y[i] - qy * q[i];
}

template <typename T>
__device__ __forceinline__ T rmax0(T x) { return x > T(0) ? x : T(0); }

// State carried from one depth step to the next (the step at depth k reads
// the upwind point k+1, the point k and the downwind point k-1).
template <typename T>
struct DeloState {
  T Iv[4];                  // Stokes vector at the upwind depth
  Point<T> Pu, P0, Pd;      // points at k+1, k, k-1
  T dSu[4], dqu[3];         // derivatives of S and q at the upwind point
  T dtau_uw, dsup, dchi_up, dchi_c;
  T dru[3];                 // derivatives of the magneto-optical r (MO only)
};

// Set-up from the three deepest points and the thermalised lower boundary
// I = B[N-1] - (B[N-2] - B[N-1]) / dtau (bezier_1D.c:93-101).  On entry
// s.Pu and s.P0 hold the points at N-1 and N-2; PNm3 is the point at N-3.
// dz1 = |z[N-2] - z[N-1]|, dz2 = |z[N-3] - z[N-2]|; B2 = (B[N-2], B[N-1]).
template <typename T>
__device__ __forceinline__ void delo_prologue(DeloState<T> &s, const Point<T> &PNm3, T dz1,
                                              T dz2, T imu, T B2a, T B2b) {
  s.Pd = PNm3;
  const T dtau0 = T(0.5) * imu * (s.Pu.chi + s.P0.chi) * dz1;
  s.Iv[0] = B2b - rdiv(B2a - B2b, dtau0);
  s.Iv[1] = s.Iv[2] = s.Iv[3] = T(0);
  s.dsup = dz1 * imu;
  const T dsdn = dz2 * imu;
  s.dchi_up = rdiv(s.P0.chi - s.Pu.chi, s.dsup);
  s.dchi_c = CentDeriv<T>(s.dsup, dsdn)(s.Pu.chi, s.P0.chi, s.Pd.chi);
  const T c2 = rmax0(s.P0.chi - (s.dsup * T(1.0 / 3.0)) * s.dchi_c);
  const T c1 = rmax0(s.Pu.chi + (s.dsup * T(1.0 / 3.0)) * s.dchi_up);
  s.dtau_uw = T(0.25) * s.dsup * (s.P0.chi + s.Pu.chi + c1 + c2);
  const T idt = rrcp(s.dtau_uw);
#pragma unroll
  for (int n = 0; n < 4; n++) s.dSu[n] = (s.P0.S[n] - s.Pu.S[n]) * idt;
#pragma unroll
  for (int n = 0; n < 3; n++) s.dqu[n] = (s.P0.q[n] - s.Pu.q[n]) * idt;
#pragma unroll
  for (int n = 0; n < 3; n++) s.dru[n] = (s.P0.r[n] - s.Pu.r[n]) * idt;
}

// The right-hand side V0 of a step and the solve Md Iv = V0 are formed in
// Wide<T>: the step's own type in FP64, FP64 in the FP32 form.  Near the
// core of a strong Zeeman line with an optically thick step, |q| -> 1 and
// Md = I + c K0 + g (K0^2 + dK0) becomes nearly singular (eigenvalues
// 1 -/+ c|q| + g|q|^2 with c ~ 2, g ~ 1).  The exact solution (I + K)^-1 S
// for a constant source only needs eps = 1 - (alpha + beta + gamma + theta),
// which float weights miss by ~1e-7; the near-singular solve amplifies that
// and float rounding of V0 to ~1e-3 of I.  So the FP32 form widens its
// (float) weights and takes eps from that identity, and builds V0 and solves
// in FP64; the FP64 form keeps the reference's eps = exp(-dtau).
template <typename T> struct Wide { using type = T; static constexpr bool widened = false; };
template <> struct Wide<float> { using type = double; static constexpr bool widened = true; };
template <> struct Wide<Dual<float>> { using type = Dual<double>; static constexpr bool widened = true; };

__device__ __forceinline__ double widen(double x) { return x; }
__device__ __forceinline__ double widen(float x) { return (double) x; }
__device__ __forceinline__ Dual<double> widen(Dual<double> x) { return x; }
__device__ __forceinline__ Dual<double> widen(Dual<float> x) {
  return Dual<double>((double) x.v, (double) x.d);
}
__device__ __forceinline__ void narrow(double x, double &y) { y = x; }
__device__ __forceinline__ void narrow(double x, float &y) { y = (float) x; }
__device__ __forceinline__ void narrow(Dual<double> x, Dual<double> &y) { y = x; }
__device__ __forceinline__ void narrow(Dual<double> x, Dual<float> &y) {
  y = Dual<float>((float) x.v, (float) x.d);
}

template <typename T>
__device__ __forceinline__ void delo_rhs_solve(DeloState<T> &s, T alpha_, T beta_, T gamma_,
                                               T theta_, T eps_, T dt03_, const T *dS0_,
                                               const T *dq0_) {
  using C = typename Wide<T>::type;
  const C alpha = widen(alpha_), beta = widen(beta_), gamma = widen(gamma_),
          theta = widen(theta_), dt03 = widen(dt03_);
  const C eps = Wide<T>::widened ? C(1) - (alpha + beta + gamma + theta) : widen(eps_);
  C Iv[4], Su[4], S0[4], dS0[4], dSu[4], qu[3], q0[3], dqu[3], dq0[3];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    Iv[i] = widen(s.Iv[i]);
    Su[i] = widen(s.Pu.S[i]);
    S0[i] = widen(s.P0.S[i]);
    dS0[i] = widen(dS0_[i]);
    dSu[i] = widen(s.dSu[i]);
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    qu[i] = widen(s.Pu.q[i]);
    q0[i] = widen(s.P0.q[i]);
    dqu[i] = widen(s.dqu[i]);
    dq0[i] = widen(dq0_[i]);
  }
  // V0 = Ma Iv + Mb Su + Mc S0 + dt03 (gamma dS0 - theta dSu) with
  // Ma = eps I - beta Ku + theta (dt03 (Ku^2 + dKu + Ku) - Ku),
  // Mb = beta I + theta (I - dt03 Ku), Mc = alpha I + gamma (I + dt03 K0)
  C KI[4], KKI[4], dKI[4], KSu[4], KS0[4], V0[4], x[4];
  kmul(qu, Iv, KI);
  kmul(qu, KI, KKI);
  kmul(dqu, Iv, dKI);
  kmul(qu, Su, KSu);
  kmul(q0, S0, KS0);
  const C ca = theta * (dt03 - C(1)) - beta, gu = theta * dt03;
  const C cb0 = beta + theta, cc0 = alpha + gamma, g0 = gamma * dt03;
#pragma unroll
  for (int i = 0; i < 4; i++)
    V0[i] = eps * Iv[i] + ca * KI[i] + gu * (KKI[i] + dKI[i]) + cb0 * Su[i] - gu * KSu[i] +
            cc0 * S0[i] + g0 * KS0[i] + dt03 * (gamma * dS0[i] - theta * dSu[i]);
  // Md = I + alpha K0 + gamma (dt03 (K0^2 + dK0 + K0) + K0)
  solve_md<C>(alpha + gamma * (C(1) + dt03), g0, q0, dq0, V0, x);
#pragma unroll
  for (int i = 0; i < 4; i++) narrow(x[i], s.Iv[i]);
}

// One step at depth k: Iv(k) from Iv(k+1).  Pn is the point at k-2 (has_pn;
// at k = 1 there is none and the downwind derivative of chi is one-sided).
// dzk = |z[k-1] - z[k]|, dzk1 = |z[k-2] - z[k-1]|.
template <typename T>
__device__ __forceinline__ void delo_step(DeloState<T> &s, const Point<T> &Pn, bool has_pn, T dzk,
                                          T dzk1, T imu) {
  const T dsdn = dzk * imu;
  T dchi_dn;
  if (has_pn) {
    const T dsdn2 = dzk1 * imu;
    dchi_dn = CentDeriv<T>(dsdn, dsdn2)(s.P0.chi, s.Pd.chi, Pn.chi);
  } else {
    dchi_dn = rdiv(s.Pd.chi - s.P0.chi, dsdn);
  }
  const T c2 = rmax0(s.P0.chi + (dsdn * T(1.0 / 3.0)) * s.dchi_c);
  const T c1 = rmax0(s.Pd.chi - (dsdn * T(1.0 / 3.0)) * dchi_dn);
  const T dtau_dw = T(0.25) * dsdn * (s.P0.chi + s.Pd.chi + c1 + c2);
  const Bezier3<T> bz = bezier3(s.dtau_uw);
  const T dt03 = s.dtau_uw * T(1.0 / 3.0);
  const T alpha = bz.alpha, beta = bz.beta, gamma = bz.gamma, theta = bz.theta, eps = bz.eps;
  T dS0[4], dq0[3];
  const CentDeriv<T> cd(s.dtau_uw, dtau_dw);
#pragma unroll
  for (int i = 0; i < 4; i++) dS0[i] = cd(s.Pu.S[i], s.P0.S[i], s.Pd.S[i]);
#pragma unroll
  for (int i = 0; i < 3; i++) dq0[i] = cd(s.Pu.q[i], s.P0.q[i], s.Pd.q[i]);
  // V0 = Ma Iv + Mb Su + Mc S0 + dt03 (gamma dS0 - theta dSu) with
  // Ma = eps I - beta Ku + theta (dt03 (Ku^2 + dKu + Ku) - Ku),
  // Mb = beta I + theta (I - dt03 Ku), Mc = alpha I + gamma (I + dt03 K0)
  delo_rhs_solve(s, alpha, beta, gamma, theta, eps, dt03, dS0, dq0);
  s.Pu = s.P0;
  s.P0 = s.Pd;
  if (has_pn) s.Pd = Pn;
#pragma unroll
  for (int n = 0; n < 4; n++) s.dSu[n] = dS0[n];
#pragma unroll
  for (int n = 0; n < 3; n++) s.dqu[n] = dq0[n];
  s.dtau_uw = dtau_dw;
  s.dsup = dsdn;
  s.dchi_up = s.dchi_c;
  s.dchi_c = dchi_dn;
}

// w3 weights of the last interval in double with their derivative in dtau
// carried by dual numbers
__device__ __forceinline__ void w3_t(double dt, double *w) { w3(dt, w); }
__device__ __forceinline__ void w3_t(float dt, float *w) {
  double wd[3];
  w3((double) dt, wd);
  w[0] = (float) wd[0]; w[1] = (float) wd[1]; w[2] = (float) wd[2];
}
template <typename R>
__device__ __forceinline__ void w3_t(Dual<R> dt, Dual<R> *w) {
  double wd[3], dw[2];
  w3d((double) dt.v, wd, dw);
  w[0] = {(R) wd[0], (R) dw[0] * dt.d};
  w[1] = {(R) wd[1], (R) dw[1] * dt.d};
  w[2] = {(R) wd[2], R(0)};
}

// The last (linear) interval, bezier_1D.c:264-299, from the state after the
// step at k = 1 (P0 = depth 0, Pu = depth 1); dz = |z[0] - z[1]|.
template <typename T>
__device__ __forceinline__ void delo_final(const DeloState<T> &s, T dz, T imu, T *x) {
  T w[3];
  const T dtl = T(0.5) * imu * (s.P0.chi + s.Pu.chi) * dz;
  w3_t(dtl, w);
  const T w0 = w[0], w1 = w[1], w1t = rdiv(w[1], dtl);
  // V0 = w0 S0 - w1 dSu + ((1 - w0) I - w1/dtau Ku) Iv;  Md = I + (w0 - w1/dtau) K0
  T KI[4], V0[4];
  kmul(s.Pu.q, s.Iv, KI);
#pragma unroll
  for (int n = 0; n < 4; n++)
    V0[n] = w0 * s.P0.S[n] - w1 * s.dSu[n] + (T(1) - w0) * s.Iv[n] - w1t * KI[n];
  const T e = w0 - w1t;
  const T q0V = s.P0.q[0] * V0[1] + s.P0.q[1] * V0[2] + s.P0.q[2] * V0[3];
  const T qq = s.P0.q[0] * s.P0.q[0] + s.P0.q[1] * s.P0.q[1] + s.P0.q[2] * s.P0.q[2];
  x[0] = rdiv(V0[0] - e * q0V, T(1) - e * e * qq);
#pragma unroll
  for (int n = 0; n < 3; n++) x[n + 1] = V0[n + 1] - e * s.P0.q[n] * x[0];
}

// ---------------------------------------------------------------------------
// The polarized step with the magneto-optical terms (MAGNETO_OPTICAL), FP64:
// K(q, r) = [[0, q^T], [q, R]] with the antisymmetric
// R = [[0, r_V, -r_U], [-r_V, 0, r_Q], [r_U, -r_Q, 0]] (rh/stokesopac.c:56-86).
// Md = I + cM K0 + gM (K0^2 + dK0) is then a general 4x4 matrix, solved by
// Gaussian elimination with partial pivoting.

// y = K(q, r) x
template <typename T>
__device__ __forceinline__ void kmul_mo(const T *q, const T *r, const T *x, T *y) {
  y[0] = q[0] * x[1] + q[1] * x[2] + q[2] * x[3];
  y[1] = q[0] * x[0] + r[2] * x[2] - r[1] * x[3];
  y[2] = q[1] * x[0] - r[2] * x[1] + r[0] * x[3];
  y[3] = q[2] * x[0] + r[1] * x[1] - r[0] * x[2];
}

// K(q, r) as a matrix
template <typename T>
__device__ __forceinline__ void kmat_mo(const T *q, const T *r, T K[4][4]) {
  K[0][0] = T(0); K[0][1] = q[0]; K[0][2] = q[1]; K[0][3] = q[2];
  K[1][0] = q[0]; K[1][1] = T(0); K[1][2] = r[2]; K[1][3] = -r[1];
  K[2][0] = q[1]; K[2][1] = -r[2]; K[2][2] = T(0); K[2][3] = r[0];
  K[3][0] = q[2]; K[3][1] = r[1]; K[3][2] = -r[0]; K[3][3] = T(0);
}

// x = A^-1 b, A 4x4 (overwritten), Gaussian elimination with partial pivoting
template <typename T>
__device__ __forceinline__ void solve4_gepp(T A[4][4], const T *b, T *x) {
  T y[4] = {b[0], b[1], b[2], b[3]};
#pragma unroll
  for (int c = 0; c < 4; c++) {
    int p = c;
#pragma unroll
    for (int r = c + 1; r < 4; r++)
      if (fabs(A[r][c]) > fabs(A[p][c])) p = r;
    if (p != c) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const T t = A[c][j]; A[c][j] = A[p][j]; A[p][j] = t;
      }
      const T t = y[c]; y[c] = y[p]; y[p] = t;
    }
    const T inv = rrcp(A[c][c]);
#pragma unroll
    for (int r = c + 1; r < 4; r++) {
      const T f = A[r][c] * inv;
#pragma unroll
      for (int j = c + 1; j < 4; j++) A[r][j] = A[r][j] - f * A[c][j];
      y[r] = y[r] - f * y[c];
    }
  }
#pragma unroll
  for (int r = 3; r >= 0; r--) {
    T acc = y[r];
#pragma unroll
    for (int j = r + 1; j < 4; j++) acc = acc - A[r][j] * x[j];
    x[r] = rdiv(acc, A[r][r]);
  }
}

// One step at depth k with the magneto-optical terms (see delo_step).
template <typename T>
__device__ __forceinline__ void delo_step_mo(DeloState<T> &s, const Point<T> &Pn, bool has_pn,
                                             T dzk, T dzk1, T imu) {
  const T dsdn = dzk * imu;
  T dchi_dn;
  if (has_pn) {
    const T dsdn2 = dzk1 * imu;
    dchi_dn = CentDeriv<T>(dsdn, dsdn2)(s.P0.chi, s.Pd.chi, Pn.chi);
  } else {
    dchi_dn = rdiv(s.Pd.chi - s.P0.chi, dsdn);
  }
  const T c2 = rmax0(s.P0.chi + (dsdn * T(1.0 / 3.0)) * s.dchi_c);
  const T c1 = rmax0(s.Pd.chi - (dsdn * T(1.0 / 3.0)) * dchi_dn);
  const T dtau_dw = T(0.25) * dsdn * (s.P0.chi + s.Pd.chi + c1 + c2);
  const Bezier3<T> bz = bezier3(s.dtau_uw);
  const T dt03 = s.dtau_uw * T(1.0 / 3.0);
  T dS0[4], dq0[3], dr0[3];
  const CentDeriv<T> cd(s.dtau_uw, dtau_dw);
#pragma unroll
  for (int i = 0; i < 4; i++) dS0[i] = cd(s.Pu.S[i], s.P0.S[i], s.Pd.S[i]);
#pragma unroll
  for (int i = 0; i < 3; i++) dq0[i] = cd(s.Pu.q[i], s.P0.q[i], s.Pd.q[i]);
#pragma unroll
  for (int i = 0; i < 3; i++) dr0[i] = cd(s.Pu.r[i], s.P0.r[i], s.Pd.r[i]);
  const T al = bz.alpha, be = bz.beta, ga = bz.gamma, th = bz.theta, ep = bz.eps;
  const T ca = th * (dt03 - T(1)) - be, gu = th * dt03, cb0 = be + th, cc0 = al + ga,
          g0 = ga * dt03, cM = al + ga * (T(1) + dt03);
  T KI[4], KKI[4], dKI[4], KSu[4], KS0[4], V0[4];
  kmul_mo(s.Pu.q, s.Pu.r, s.Iv, KI);
  kmul_mo(s.Pu.q, s.Pu.r, KI, KKI);
  kmul_mo(s.dqu, s.dru, s.Iv, dKI);
  kmul_mo(s.Pu.q, s.Pu.r, s.Pu.S, KSu);
  kmul_mo(s.P0.q, s.P0.r, s.P0.S, KS0);
#pragma unroll
  for (int i = 0; i < 4; i++)
    V0[i] = ep * s.Iv[i] + ca * KI[i] + gu * (KKI[i] + dKI[i]) + cb0 * s.Pu.S[i] - gu * KSu[i] +
            cc0 * s.P0.S[i] + g0 * KS0[i] + dt03 * (ga * dS0[i] - th * s.dSu[i]);
  T K0[4][4], dK0[4][4], Md[4][4];
  kmat_mo(s.P0.q, s.P0.r, K0);
  kmat_mo(dq0, dr0, dK0);
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      T kk = T(0);
#pragma unroll
      for (int m = 0; m < 4; m++) kk += K0[i][m] * K0[m][j];
      Md[i][j] = (i == j ? T(1) : T(0)) + cM * K0[i][j] + g0 * (kk + dK0[i][j]);
    }
  solve4_gepp(Md, V0, s.Iv);
  s.Pu = s.P0;
  s.P0 = s.Pd;
  if (has_pn) s.Pd = Pn;
#pragma unroll
  for (int n = 0; n < 4; n++) s.dSu[n] = dS0[n];
#pragma unroll
  for (int n = 0; n < 3; n++) { s.dqu[n] = dq0[n]; s.dru[n] = dr0[n]; }
  s.dtau_uw = dtau_dw;
  s.dsup = dsdn;
  s.dchi_up = s.dchi_c;
  s.dchi_c = dchi_dn;
}

// The last interval with the magneto-optical terms (see delo_final):
// Md = I + (w0 - w1/dtau) K0, V0 = w0 S0 - w1 dSu + ((1 - w0) I - w1/dtau Ku) Iv.
template <typename T>
__device__ __forceinline__ void delo_final_mo(const DeloState<T> &s, T dz, T imu, T *x) {
  T w[3];
  const T dtl = T(0.5) * imu * (s.P0.chi + s.Pu.chi) * dz;
  w3_t(dtl, w);
  const T w0 = w[0], w1 = w[1], w1t = rdiv(w[1], dtl);
  T KI[4], V0[4], Md[4][4];
  kmul_mo(s.Pu.q, s.Pu.r, s.Iv, KI);
#pragma unroll
  for (int n = 0; n < 4; n++)
    V0[n] = w0 * s.P0.S[n] - w1 * s.dSu[n] + (T(1) - w0) * s.Iv[n] - w1t * KI[n];
  const T e = w0 - w1t;
  kmat_mo(s.P0.q, s.P0.r, Md);
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) Md[i][j] = (i == j ? T(1) : T(0)) + e * Md[i][j];
  solve4_gepp(Md, V0, x);
}

}  // namespace rhg
