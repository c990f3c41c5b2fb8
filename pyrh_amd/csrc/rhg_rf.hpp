// Response functions of the emergent Stokes vector to the atmosphere,
// dStokes(lambda)/d p(k) for p in {T, v_LOS, B, gamma, chi, v_mic} at every
// depth k, for the polarized (DELO-Bezier) wavelengths.
//
// Four parts:
//  1. Tangents of the column state.  The equilibria of one depth depend
//     only on that depth's inputs, so one central difference with every
//     depth perturbed at once gives d(field)/d p(k) at all depths
//     (k_rf_perturb + k_state twice, FP64); the same for the continuum and
//     for chi at the reference wavelength.
//  2. k_rf_forward: the synthesis of k_formal_stokes that also keeps, per
//     ray, the 13 carried scalars of the recursion entering each depth step
//     (Stokes vector, derivatives of S and q, optical-depth interval, chi
//     derivative); not the points.
//  3. k_rf_reverse: one reverse depth sweep per ray for the four outputs
//     I, Q, U, V at once, with the hand-written vector-Jacobian product of
//     each step (delo_step_vjp, rhg_adjoint.hpp).  The points (opacity,
//     source functions, reduced propagation matrix) are formed again as the
//     sweep reaches them.  Each point's adjoint is stored once, when the
//     last step reading it is done, with those of the path lengths |dz| and
//     of the lower boundary's Planck values.
//  4. k_rf_chain: per (ray, depth), the point's derivative along each
//     parameter's state tangent (form_point on dual numbers: line opacity,
//     Zeeman-split Voigt/Faraday-Voigt profiles, continuum), contracted with
//     the point adjoints, plus the path-length terms (tau500 scale: dz(j)
//     depends on chi_ref at j-1 and j) and the boundary terms.
// The recursion, its adjoint and the points are FP64 in both forms; the FP32
// form evaluates the Voigt functions (and their derivatives) in float.
#pragma once

#include "rhg_adjoint.hpp"

namespace rhg {

template <typename T>
__device__ __forceinline__ T &point_ref(Point<T> &p, int m) {
  switch (m) {
    case 0: return p.chi;
    case 1: return p.S[0];
    case 2: return p.S[1];
    case 3: return p.S[2];
    case 4: return p.S[3];
    case 5: return p.q[0];
    case 6: return p.q[1];
    default: return p.q[2];
  }
}

template <typename T>
__device__ __forceinline__ T point_get(const Point<T> &p, int m) {
  switch (m) {
    case 0: return p.chi;
    case 1: return p.S[0];
    case 2: return p.S[1];
    case 3: return p.S[2];
    case 4: return p.S[3];
    case 5: return p.q[0];
    case 6: return p.q[1];
    default: return p.q[2];
  }
}

// ---------------------------------------------------------------------------
// 2. forward pass storing the carried recursion state

// carry[((t * Nmax + k) * NCARRY_STEP + c) * ncol + column] (t = entry of the
// wavelength list): the carried state entering the step at depth k
// (k = 1 .. N-2); k = 0 holds the state entering the last interval.  The
// points (opacity, source functions, K') are not stored: the reverse sweep
// forms them again.
#define RF_SETUP()                                                                       \
  const DevModel &M = *Mp;                                                               \
  int cb, t;                                                                             \
  xcd_tile(blockIdx.x, nlist, cb, t);                                                    \
  const int ic = cb * 64 + threadIdx.x;                                                  \
  if (ic >= ncol) return;                                                                \
  const int io = io_list[t];                                                             \
  const int N = ndep[ic];                                                                \
  const int ls = out_ls[io];                                                             \
  const int moving = moving_arr[ic];                                                     \
  const int64_t fs = (int64_t) Nmax * ncol;                                              \
  const double *st0 = state + ic;                                                        \
  const double *dzc = st0 + (int64_t) RHG_CF_DZ * fs;                                    \
  const double *Tc = st0 + (int64_t) RHG_CF_T * fs;                                      \
  const double *lc = M.lamconst + (int64_t) ls * RHG_LC_NCONST;                          \
  (void) lc; (void) Tc; (void) moving;
#define DZ(kk) dzc[(int64_t) (kk) * ncol]
#define CAR(kk, c) car[((((int64_t) t * Nmax + (kk)) * NCARRY_STEP + (c)) * ncol) + ic]
// the point at depth jj: FP64, with the Voigt functions in R
#define RF_POINT(jj, P)                                                                  \
  form_point_v<R>(M, ls, has_line, st0 + (int64_t) (jj) * ncol, fs, mu_d, moving,        \
                  cc[(int64_t) (jj) * ncol], ce[(int64_t) (jj) * ncol], P)

template <typename R>
__global__ void __launch_bounds__(64)
k_rf_forward(const DevModel *__restrict__ Mp, const int32_t *__restrict__ ndep,
             const int32_t *__restrict__ moving_arr, int ncol, int Nmax,
             const double *__restrict__ state, const int32_t *__restrict__ out_ls,
             const int32_t *__restrict__ io_list, int nlist, int nout,
             const R *__restrict__ cchi, const R *__restrict__ ceta, double mu_d,
             double *__restrict__ stokes, double *__restrict__ car) {
  RF_SETUP();
  const R *cc = cchi + (int64_t) io * Nmax * ncol + ic;
  const R *ce = ceta + (int64_t) io * Nmax * ncol + ic;
  const int has_line = ldc(M.lam_line_ptr)[ls + 1] > ldc(M.lam_line_ptr)[ls];
  const double imu = 1.0 / mu_d;
  const double B2a = planck(Tc[(int64_t) (N - 2) * ncol], lc[RHG_LC_HC_KLA], lc[RHG_LC_TWOHNU3_C2]);
  const double B2b = planck(Tc[(int64_t) (N - 1) * ncol], lc[RHG_LC_HC_KLA], lc[RHG_LC_TWOHNU3_C2]);
  DeloState<double> s;
  Point<double> Pn;
  double cv[NCARRY_STEP];
  for (int j = N - 1; j >= -1; j--) {
    if (j >= 0) RF_POINT(j, Pn);
    if (j == N - 1) {
      s.Pu = Pn;
    } else if (j == N - 2) {
      s.P0 = Pn;
    } else if (j == N - 3) {
      delo_prologue<double>(s, Pn, DZ(N - 1), DZ(N - 2), imu, B2a, B2b);
    } else {
      const int k = j + 2;
      carry_to_array(s, cv);
#pragma unroll
      for (int c = 0; c < NCARRY_STEP; c++) CAR(k, c) = cv[c];
      delo_step<double>(s, Pn, j >= 0, DZ(k), j >= 0 ? DZ(k - 1) : 0.0, imu);
    }
  }
  carry_to_array(s, cv);
#pragma unroll
  for (int c = 0; c < NCARRY_STEP; c++) CAR(0, c) = cv[c];
  double x[4];
  delo_final<double>(s, DZ(1), imu, x);
  double *out = stokes + (int64_t) ic * 4 * nout + io;
#pragma unroll
  for (int i = 0; i < 4; i++) out[(int64_t) i * nout] = x[i];
}

// ---------------------------------------------------------------------------
// 3. reverse sweep

// adjoint layout (t = list entry, o = output I, Q, U, V):
//   aP[(((t * 4 + o) * Nmax + k) * 8 + m) * ncol + column]   point k (chi, S, q)
//   aZ[((t * 4 + o) * Nmax + k) * ncol + column]     (k = 1 .. N-1: |dz(k)|)
//   aB[((t * 4 + o) * 2 + i) * ncol + column]          (i = 0: B(N-2), 1: B(N-1))
#define AP(o, kk, m) aP[(((((int64_t) t * 4 + (o)) * Nmax + (kk)) * 8 + (m)) * ncol) + ic]
#define AZ(o, kk) aZ[((((int64_t) t * 4 + (o)) * Nmax + (kk)) * ncol) + ic]
#define AB(o, i) aB[((((int64_t) t * 4 + (o)) * 2 + (i)) * ncol) + ic]

__device__ __forceinline__ void carry_from_array(const double *c, StepCarry &s) {
#pragma unroll
  for (int i = 0; i < 4; i++) s.Iv[i] = c[i];
#pragma unroll
  for (int i = 0; i < 4; i++) s.dSu[i] = c[4 + i];
#pragma unroll
  for (int i = 0; i < 3; i++) s.dqu[i] = c[8 + i];
  s.dtau_uw = c[11];
  s.dchi_c = c[12];
}

// The point adjoints of a sweep live in LDS, in a ring of four depth slots
// (depth j in slot j & 3) per output and component: [slot][o][m][lane].
// Each lane touches only its own entries.  add() takes the slots of the
// VJP routines (0 = Pu, 1 = P0, 2 = Pd, 3 = Pn) relative to the depth kp1 of
// Pu.
template <typename A>
struct PbarRing {
  A *base;
  int lane, kp1;
  __device__ __forceinline__ A &at(int depth, int o, int m) const {
    return base[((((depth & 3) * 4 + o) * 8) + m) * 64 + lane];
  }
  __device__ __forceinline__ void add(int slot, int o, int m, double v) const {
    at(kp1 - slot, o, m) += A(v);
  }
};

// One reverse sweep per (column, wavelength) for the four outputs at once:
// the last interval, each depth step and the prologue by their hand-written
// vector-Jacobian products (rhg_adjoint.hpp).  The points are formed again,
// one per step (the point k+1 entering the window of step k); the carried
// state entering each step comes from the forward pass.  A point's adjoint
// is complete after the step at which it is the second point downwind
// (k = j + 2) and is stored then, once; a path length's after the step
// downwind of it.
template <typename R>
__global__ void __launch_bounds__(64)
k_rf_reverse(const DevModel *__restrict__ Mp, const int32_t *__restrict__ ndep,
             const int32_t *__restrict__ moving_arr, int ncol, int Nmax,
             const double *__restrict__ state, const int32_t *__restrict__ out_ls,
             const int32_t *__restrict__ io_list, int nlist, const R *__restrict__ cchi,
             const R *__restrict__ ceta, double mu_d, const double *__restrict__ car,
             R *__restrict__ aP, R *__restrict__ aZ, R *__restrict__ aB) {
  __shared__ R ring_lds[4 * 4 * 8 * 64];
  RF_SETUP();
  const R *cc = cchi + (int64_t) io * Nmax * ncol + ic;
  const R *ce = ceta + (int64_t) io * Nmax * ncol + ic;
  const int has_line = ldc(M.lam_line_ptr)[ls + 1] > ldc(M.lam_line_ptr)[ls];
  const double imu = 1.0 / mu_d;
  PbarRing<R> ring{ring_lds, (int) threadIdx.x, 1};
  for (int d = 0; d < 4; d++)
#pragma unroll
    for (int o = 0; o < 4; o++)
#pragma unroll
      for (int m = 0; m < 8; m++) ring.at(d, o, m) = R(0);

  // window of points: Pw[0..3] = depths k+1, k, k-1, k-2 at step k (shifted
  // before each step); path-length adjoints |dz(k)|, |dz(k-1)|
  Point<double> Pw[4];
  R zb[4][2];
  StepCarry Cb[4];
#pragma unroll
  for (int o = 0; o < 4; o++) zb[o][0] = zb[o][1] = R(0);

  // --- last interval: Pu = depth 1, P0 = depth 0
  RF_POINT(1, Pw[0]);
  RF_POINT(0, Pw[1]);
  {
    double cv[NCARRY_STEP];
#pragma unroll
    for (int c = 0; c < NCARRY_STEP; c++) cv[c] = CAR(0, c);
    StepCarry C0;
    carry_from_array(cv, C0);
    const double xb[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    R dzb[4] = {R(0), R(0), R(0), R(0)};
    ring.kp1 = 1;
    delo_final_vjp<4>(C0, Pw[0], Pw[1], DZ(1), imu, xb, Cb, ring, dzb);
#pragma unroll
    for (int o = 0; o < 4; o++) zb[o][0] = dzb[o];
  }
#pragma unroll
  for (int o = 0; o < 4; o++) AZ(o, 0) = R(0);

  // --- steps k = 1 .. N-2, the reverse of the forward order
  for (int k = 1; k <= N - 2; k++) {
    Pw[3] = Pw[2];
    Pw[2] = Pw[1];
    Pw[1] = Pw[0];
    RF_POINT(k + 1, Pw[0]);
    double cv[NCARRY_STEP];
#pragma unroll
    for (int c = 0; c < NCARRY_STEP; c++) cv[c] = CAR(k, c);
    StepCarry C;
    carry_from_array(cv, C);
    const bool has_pn = k >= 2;
    R dzb[4], dz1b[4];
#pragma unroll
    for (int o = 0; o < 4; o++) { dzb[o] = zb[o][0]; dz1b[o] = zb[o][1]; }
    ring.kp1 = k + 1;
    delo_step_vjp<4>(C, Pw[0], Pw[1], Pw[2], Pw[3], has_pn, DZ(k), has_pn ? DZ(k - 1) : 0.0,
                     imu, Cb, ring, dzb, dz1b);
    // point k-2 and |dz(k-1)| are complete; the slot of k-2 is k+2's next
    if (has_pn) {
#pragma unroll
      for (int o = 0; o < 4; o++) {
#pragma unroll
        for (int m = 0; m < 8; m++) {
          AP(o, k - 2, m) = ring.at(k - 2, o, m);
          ring.at(k - 2, o, m) = R(0);
        }
        AZ(o, k - 1) = dz1b[o];
      }
    }
#pragma unroll
    for (int o = 0; o < 4; o++) {
      zb[o][1] = dzb[o];
      zb[o][0] = R(0);
    }
  }
  // Pw[0..2] = depths N-1, N-2, N-3; zb = |dz(N-1)|, |dz(N-2)|

  // --- prologue
  {
    const double B2a = planck(Tc[(int64_t) (N - 2) * ncol], lc[RHG_LC_HC_KLA], lc[RHG_LC_TWOHNU3_C2]);
    const double B2b = planck(Tc[(int64_t) (N - 1) * ncol], lc[RHG_LC_HC_KLA], lc[RHG_LC_TWOHNU3_C2]);
    R dz1b[4], dz2b[4], bab[4], bbb[4];
#pragma unroll
    for (int o = 0; o < 4; o++) {
      dz1b[o] = zb[o][0];
      dz2b[o] = zb[o][1];
      bab[o] = bbb[o] = R(0);
    }
    ring.kp1 = N - 1;
    delo_prologue_vjp<4>(Pw[0], Pw[1], Pw[2], DZ(N - 1), DZ(N - 2), imu, B2a, B2b, Cb, ring,
                         dz1b, dz2b, bab, bbb);
#pragma unroll
    for (int o = 0; o < 4; o++) {
      AZ(o, N - 1) = dz1b[o];
      AZ(o, N - 2) = dz2b[o];
      AB(o, 0) = bab[o];
      AB(o, 1) = bbb[o];
#pragma unroll
      for (int m = 0; m < 8; m++) {
        AP(o, N - 1, m) = ring.at(N - 1, o, m);
        AP(o, N - 2, m) = ring.at(N - 2, o, m);
        AP(o, N - 3, m) = ring.at(N - 3, o, m);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 4. chain to the atmosphere parameters

// rf[((((i * Nmax + k) * 4 + o) * nout + io) * ldr + column] for the i-th
// requested parameter pl.p[i] (i < npar <= NP); chiref and dchiref are
// [column][depth] (as k_geometry writes chi_ref).  One pass of form_point on
// NP-tangent dual numbers per depth gives d(point)/dp for every requested
// parameter (the Humlicek functions by their complex derivative).
// The FP32 form's three-tangent chain is held to two waves per SIMD (it
// needs 292 registers at one; two cost 36 spilled to scratch and run 1024
// columns of 64-256 depths under 400 Zeeman lines 1.6x faster); the others
// keep one.
template <typename R, int NP>
__global__ void __launch_bounds__(64, (NP <= 3 && sizeof(R) == 4 ? 2 : 1))
k_rf_chain(const DevModel *__restrict__ Mp, const int32_t *__restrict__ ndep,
           const int32_t *__restrict__ moving_arr, int ncol, int Nmax,
           const double *__restrict__ state, const double *__restrict__ dstate,
           const double *__restrict__ chiref, const double *__restrict__ dchiref,
           const R *__restrict__ cchi, const R *__restrict__ ceta,
           const R *__restrict__ dcchi, const R *__restrict__ dceta,
           const int32_t *__restrict__ out_ls, const int32_t *__restrict__ io_list, int nlist,
           int nout, double mu_d, int atm_scale, RfParamList pl, const R *__restrict__ aP,
           const R *__restrict__ aZ, const R *__restrict__ aB, R *__restrict__ rf, int64_t ldr) {
  RF_SETUP();
  using D = DualN<double, NP>;
  using DV = DualN<R, NP>;
  const int has_line = ldc(M.lam_line_ptr)[ls + 1] > ldc(M.lam_line_ptr)[ls];
  const D muD = D(mu_d);
  const int64_t npts = (int64_t) Nmax * ncol;
  const int64_t tan_stride = (int64_t) state_nfield(M.ncont, M.nline) * npts;
  int iT = -1;
#pragma unroll
  for (int i = 0; i < NP; i++)
    if (i < pl.npar && pl.p[i] == RHG_RF_T) iT = i;
  for (int k = 0; k < N; k++) {
    const double *stk = st0 + (int64_t) k * ncol;
    const int64_t ck = ((int64_t) io * Nmax + k) * ncol + ic;
    const double *dst[NP];
#pragma unroll
    for (int i = 0; i < NP; i++)
      dst[i] = i < pl.npar ? dstate + pl.p[i] * tan_stride + (int64_t) k * ncol + ic : nullptr;
    auto F = [stk, &dst, fs, &pl](int f) {
      D x(stk[f * fs]);
#pragma unroll
      for (int i = 0; i < NP; i++) x.d[i] = i < pl.npar ? dst[i][f * fs] : 0.0;
      return x;
    };
    D cch((double) cchi[ck]), cet((double) ceta[ck]);
    if (iT >= 0) {
#pragma unroll
      for (int i = 0; i < NP; i++)
        if (i == iT) { cch.d[i] = (double) dcchi[ck]; cet.d[i] = (double) dceta[ck]; }
    }
    Point<D> P;
    form_point_f<D, DV>(M, ls, has_line, F, muD, moving, cch, cet, P);
    double res[NP][4];
#pragma unroll
    for (int o = 0; o < 4; o++) {
      double ap[8];
#pragma unroll
      for (int m = 0; m < 8; m++) ap[m] = (double) AP(o, k, m);
#pragma unroll
      for (int i = 0; i < NP; i++) {
        double acc = ap[0] * P.chi.d[i];
#pragma unroll
        for (int n = 0; n < 4; n++) acc += ap[1 + n] * P.S[n].d[i];
#pragma unroll
        for (int n = 0; n < 3; n++) acc += ap[5 + n] * P.q[n].d[i];
        res[i][o] = acc;
      }
    }
    if (iT >= 0) {
      double extra[4] = {0.0, 0.0, 0.0, 0.0};
      // path lengths on the tau500 scale: |dz(j)| = 2 (tau_j - tau_{j-1}) /
      // (chi_ref(j-1) + chi_ref(j)), j = k and k + 1
      if (atm_scale == 0) {
        const double dchi = dchiref[(int64_t) ic * Nmax + k];
#pragma unroll
        for (int jj = 0; jj < 2; jj++) {
          const int j = k + jj;
          if (j >= 1 && j <= N - 1) {
            const double den = chiref[(int64_t) ic * Nmax + j - 1] + chiref[(int64_t) ic * Nmax + j];
            const double dzdT = -DZ(j) / den * dchi;
#pragma unroll
            for (int o = 0; o < 4; o++) extra[o] += (double) AZ(o, j) * dzdT;
          }
        }
      }
      // the thermalised lower boundary's Planck values
      if (k >= N - 2) {
        const double T = Tc[(int64_t) k * ncol], x = lc[RHG_LC_HC_KLA] / T;
        double dB = 0.0;
        if (x <= MAX_EXPONENT) {
          const double e = exp(x), B = lc[RHG_LC_TWOHNU3_C2] / (e - 1.0);
          dB = B * x * e / (e - 1.0) / T;
        }
        const double dTdp = dstate[RHG_RF_T * tan_stride + (int64_t) RHG_CF_T * fs + (int64_t) k * ncol + ic];
#pragma unroll
        for (int o = 0; o < 4; o++) extra[o] += (double) AB(o, k - (N - 2)) * (dB * dTdp);
      }
#pragma unroll
      for (int i = 0; i < NP; i++)
        if (i == iT)
#pragma unroll
          for (int o = 0; o < 4; o++) res[i][o] += extra[o];
    }
#pragma unroll
    for (int i = 0; i < NP; i++)
      if (i < pl.npar)
#pragma unroll
        for (int o = 0; o < 4; o++)
          rf[((((int64_t) i * Nmax + k) * 4 + o) * nout + io) * ldr + ic] = (R) res[i][o];
  }
}

#undef RF_SETUP
#undef RF_POINT
#undef DZ
#undef CAR
#undef AP
#undef AZ
#undef AB

#define RHG_RF_INST(R)                                                                          \
  template __global__ void k_rf_forward<R>(const DevModel *, const int32_t *, const int32_t *,  \
                                           int, int, const double *, const int32_t *,          \
                                           const int32_t *, int, int, const R *, const R *,    \
                                           double, double *, double *);                        \
  template __global__ void k_rf_reverse<R>(const DevModel *, const int32_t *, const int32_t *,  \
                                           int, int, const double *, const int32_t *,          \
                                           const int32_t *, int, const R *, const R *, double, \
                                           const double *, R *, R *, R *);                     \
  template __global__ void k_rf_chain<R, 3>(const DevModel *, const int32_t *, const int32_t *, \
                                            int, int, const double *, const double *,          \
                                            const double *, const double *, const R *,         \
                                            const R *, const R *, const R *, const int32_t *,  \
                                            const int32_t *, int, int, double, int,            \
                                            RfParamList, const R *, const R *, const R *, R *, \
                                            int64_t);                                          \
  template __global__ void k_rf_chain<R, 6>(const DevModel *, const int32_t *, const int32_t *, \
                                            int, int, const double *, const double *,          \
                                            const double *, const double *, const R *,         \
                                            const R *, const R *, const R *, const int32_t *,  \
                                            const int32_t *, int, int, double, int,            \
                                            RfParamList, const R *, const R *, const R *, R *, \
                                            int64_t);

}  // namespace rhg
