// Response functions of the emergent Stokes vector to the atmosphere,
// dStokes(lambda)/d p(k) for p in {T, v_LOS, B, gamma, chi, v_mic} at every
// depth k, for the polarized (DELO-Bezier) wavelengths.
//
// Three parts:
//  1. Tangents of the column state.  The equilibria of one depth depend
//     only on that depth's inputs, so one central difference with every
//     depth perturbed at once gives d(field)/d p(k) at all depths
//     (k_rf_perturb + k_state twice, FP64); the same for the continuum and
//     for chi at the reference wavelength.
//  2. k_rf_forward: the synthesis of k_formal_stokes that also keeps, per
//     ray, every point (chi, S, q) and the carried part of the recursion
//     state entering each depth step (the tape).
//  3. k_rf_reverse: one reverse depth sweep per ray carrying the adjoint of
//     the 15 carried scalars for the four outputs I, Q, U, V.  Each step's
//     partial derivatives come from evaluating delo_step on dual numbers
//     with one seeded input at a time, the same code the synthesis runs.
//     Adjoints of the points, of the path lengths |dz| and of the lower
//     boundary's Planck values accumulate per depth.
//  4. k_rf_chain: per (ray, depth), the point's derivative along each
//     parameter's state tangent (form_point on dual numbers: line opacity,
//     Zeeman-split Voigt/Faraday-Voigt profiles, continuum), contracted with
//     the point adjoints, plus the path-length terms (tau500 scale: dz(j)
//     depends on chi_ref at j-1 and j) and the boundary terms.
#include "rhg_delo.hpp"

namespace rhg {

// the carried (non-point) scalars of DeloState (NCARRY of them), in a fixed
// order

template <typename T>
__device__ __forceinline__ T &carry_ref(DeloState<T> &s, int c) {
  // c is a compile-time index wherever this is used (unrolled loops)
  switch (c) {
    case 0: return s.Iv[0];
    case 1: return s.Iv[1];
    case 2: return s.Iv[2];
    case 3: return s.Iv[3];
    case 4: return s.dSu[0];
    case 5: return s.dSu[1];
    case 6: return s.dSu[2];
    case 7: return s.dSu[3];
    case 8: return s.dqu[0];
    case 9: return s.dqu[1];
    case 10: return s.dqu[2];
    case 11: return s.dtau_uw;
    case 12: return s.dsup;
    case 13: return s.dchi_up;
    default: return s.dchi_c;
  }
}

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

// plain -> dual with the derivative of one selected scalar set to 1 (none
// if seed is outside 0..7)
template <typename R>
__device__ __forceinline__ void seed_point(const Point<R> &p, Point<Dual<R>> &d, int seed) {
#pragma unroll
  for (int m = 0; m < 8; m++) point_ref(d, m) = Dual<R>(point_get(p, m), m == seed ? R(1) : R(0));
}

// ---------------------------------------------------------------------------
// 1. state tangents

// out = atmos with row `row` shifted by delta (every column and depth)
__global__ void k_rf_perturb(const double *__restrict__ atmos, double *__restrict__ out,
                             int64_t n, int Nmax, int row, double delta) {
  const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = (int) ((i / Nmax) % 9);
  out[i] = atmos[i] + (r == row ? delta : 0.0);
}

// out = (a - b) * scale, elementwise, in the output type
template <typename O>
__global__ void k_rf_diff(const double *__restrict__ a, const double *__restrict__ b, int64_t n,
                          double scale, O *__restrict__ out) {
  const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = (O) ((a[i] - b[i]) * scale);
}
template __global__ void k_rf_diff<double>(const double *, const double *, int64_t, double, double *);
template __global__ void k_rf_diff<float>(const double *, const double *, int64_t, double, float *);

// ---------------------------------------------------------------------------
// 2. forward pass with the tape

// tape layout (t = entry of the wavelength list):
//   pts[((t * Nmax + k) * 8 + m) * ncol + column]     point at depth k
//   car[((t * Nmax + k) * NCARRY + c) * ncol + column] carried state entering
//     the step at depth k (k = 1 .. N-2); k = 0 holds the state entering the
//     last interval
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
#define PTS(kk, m) pts[((((int64_t) t * Nmax + (kk)) * 8 + (m)) * ncol) + ic]
#define CAR(kk, c) car[((((int64_t) t * Nmax + (kk)) * NCARRY + (c)) * ncol) + ic]

template <typename R>
__global__ void __launch_bounds__(64)
k_rf_forward(const DevModel *__restrict__ Mp, const int32_t *__restrict__ ndep,
             const int32_t *__restrict__ moving_arr, int ncol, int Nmax,
             const double *__restrict__ state, const int32_t *__restrict__ out_ls,
             const int32_t *__restrict__ io_list, int nlist, int nout,
             const R *__restrict__ cchi, const R *__restrict__ ceta, double mu_d,
             double *__restrict__ stokes, R *__restrict__ pts, R *__restrict__ car) {
  RF_SETUP();
  const R *cc = cchi + (int64_t) io * Nmax * ncol + ic;
  const R *ce = ceta + (int64_t) io * Nmax * ncol + ic;
  const int has_line = M.lam_line_ptr[ls + 1] > M.lam_line_ptr[ls];
  const R mu = (R) mu_d, imu = (R) (1.0 / mu_d);
  const R B2a = (R) planck(Tc[(int64_t) (N - 2) * ncol], lc[RHG_LC_HC_KLA], lc[RHG_LC_TWOHNU3_C2]);
  const R B2b = (R) planck(Tc[(int64_t) (N - 1) * ncol], lc[RHG_LC_HC_KLA], lc[RHG_LC_TWOHNU3_C2]);
  DeloState<R> s;
  Point<R> Pn;
  for (int j = N - 1; j >= -1; j--) {
    if (j >= 0) {
      const double *stj = st0 + (int64_t) j * ncol;
      auto F = [stj, fs](int f) { return (R) stj[f * fs]; };
      form_point_f<R>(M, ls, has_line, F, mu, moving, cc[(int64_t) j * ncol],
                      ce[(int64_t) j * ncol], Pn);
#pragma unroll
      for (int m = 0; m < 8; m++) PTS(j, m) = point_ref(Pn, m);
    }
    if (j == N - 1) {
      s.Pu = Pn;
    } else if (j == N - 2) {
      s.P0 = Pn;
    } else if (j == N - 3) {
      delo_prologue<R>(s, Pn, (R) DZ(N - 1), (R) DZ(N - 2), imu, B2a, B2b);
    } else {
      const int k = j + 2;
#pragma unroll
      for (int c = 0; c < NCARRY; c++) CAR(k, c) = carry_ref(s, c);
      delo_step<R>(s, Pn, j >= 0, (R) DZ(k), (R) (j >= 0 ? DZ(k - 1) : 0.0), imu);
    }
  }
#pragma unroll
  for (int c = 0; c < NCARRY; c++) CAR(0, c) = carry_ref(s, c);
  R x[4];
  delo_final<R>(s, (R) DZ(1), imu, x);
  double *out = stokes + (int64_t) ic * 4 * nout + io;
#pragma unroll
  for (int i = 0; i < 4; i++) out[(int64_t) i * nout] = (double) x[i];
}

// ---------------------------------------------------------------------------
// 3. reverse sweep

// adjoint layout (t = list entry, o = output I, Q, U, V):
//   aP[(((t * 4 + o) * Nmax + k) * 8 + m) * ncol + column]
//   aZ[((t * 4 + o) * Nmax + k) * ncol + column]     (k = 1 .. N-1: |dz(k)|)
//   aB[((t * 4 + o) * 2 + i) * ncol + column]          (i = 0: B(N-2), 1: B(N-1))
#define AP(o, kk, m) aP[(((((int64_t) t * 4 + (o)) * Nmax + (kk)) * 8 + (m)) * ncol) + ic]
#define AZ(o, kk) aZ[((((int64_t) t * 4 + (o)) * Nmax + (kk)) * ncol) + ic]
#define AB(o, i) aB[((((int64_t) t * 4 + (o)) * 2 + (i)) * ncol) + ic]

template <typename R>
__device__ __forceinline__ void load_point(const R *__restrict__ pts, int64_t t, int Nmax, int k,
                                           int ncol, int ic, Point<R> &p) {
#pragma unroll
  for (int m = 0; m < 8; m++) point_ref(p, m) = pts[((((int64_t) t * Nmax + k) * 8 + m) * ncol) + ic];
}

template <typename R>
__global__ void __launch_bounds__(64)
k_rf_reverse(const DevModel *__restrict__ Mp, const int32_t *__restrict__ ndep,
             const int32_t *__restrict__ moving_arr, int ncol, int Nmax,
             const double *__restrict__ state, const int32_t *__restrict__ out_ls,
             const int32_t *__restrict__ io_list, int nlist, double mu_d,
             const R *__restrict__ pts, const R *__restrict__ car, R *__restrict__ aP,
             R *__restrict__ aZ, R *__restrict__ aB) {
  RF_SETUP();
  using D = Dual<R>;
  const R imu = (R) (1.0 / mu_d);
  const D imuD = D(imu);
  __shared__ R sbar_new[4][NCARRY][64];
  R sbar[4][NCARRY];

  // zero this ray's accumulators
  for (int k = 0; k < N; k++)
#pragma unroll
    for (int o = 0; o < 4; o++) {
#pragma unroll
      for (int m = 0; m < 8; m++) AP(o, k, m) = R(0);
      AZ(o, k) = R(0);
    }

  // --- last interval: x = delo_final(s, |dz(1)|): seeds are the 15 carried
  // scalars, the points at depths 1 (Pu) and 0 (P0), and |dz(1)|
  {
    DeloState<R> s;
    load_point(pts, t, Nmax, 1, ncol, ic, s.Pu);
    load_point(pts, t, Nmax, 0, ncol, ic, s.P0);
    s.Pd = s.P0;
#pragma unroll
    for (int c = 0; c < NCARRY; c++) carry_ref(s, c) = CAR(0, c);
    const R dz1 = (R) DZ(1);
    for (int seed = 0; seed < NCARRY + 17; seed++) {
      DeloState<D> d;
      seed_point(s.Pu, d.Pu, seed - NCARRY);
      seed_point(s.P0, d.P0, seed - NCARRY - 8);
      d.Pd = d.P0;
#pragma unroll
      for (int c = 0; c < NCARRY; c++) carry_ref(d, c) = D(carry_ref(s, c), c == seed ? R(1) : R(0));
      const D dzD(dz1, seed == NCARRY + 16 ? R(1) : R(0));
      D x[4];
      delo_final<D>(d, dzD, imuD, x);
      if (seed < NCARRY) {
#pragma unroll
        for (int o = 0; o < 4; o++) sbar_new[o][seed][threadIdx.x] = x[o].d;
      } else if (seed < NCARRY + 8) {
#pragma unroll
        for (int o = 0; o < 4; o++) AP(o, 1, seed - NCARRY) += x[o].d;
      } else if (seed < NCARRY + 16) {
#pragma unroll
        for (int o = 0; o < 4; o++) AP(o, 0, seed - NCARRY - 8) += x[o].d;
      } else {
#pragma unroll
        for (int o = 0; o < 4; o++) AZ(o, 1) += x[o].d;
      }
    }
#pragma unroll
    for (int o = 0; o < 4; o++)
#pragma unroll
      for (int c = 0; c < NCARRY; c++) sbar[o][c] = sbar_new[o][c][threadIdx.x];
  }

  // --- steps k = 1 .. N-2 (the reverse of the forward order).  Step k reads
  // the carried state CAR(k), the points k+1 (Pu), k (P0), k-1 (Pd) and,
  // for k >= 2, k-2 (Pn), and |dz(k)|, |dz(k-1)|.  Seeds: 15 carried, 8 x 4
  // points, 2 path lengths.
  for (int k = 1; k <= N - 2; k++) {
    DeloState<R> s;
    Point<R> Pn;
    load_point(pts, t, Nmax, k + 1, ncol, ic, s.Pu);
    load_point(pts, t, Nmax, k, ncol, ic, s.P0);
    load_point(pts, t, Nmax, k - 1, ncol, ic, s.Pd);
    const bool has_pn = k >= 2;
    if (has_pn) load_point(pts, t, Nmax, k - 2, ncol, ic, Pn);
    else Pn = s.Pd;
#pragma unroll
    for (int c = 0; c < NCARRY; c++) carry_ref(s, c) = CAR(k, c);
    const R dzk = (R) DZ(k), dzk1 = (R) (has_pn ? DZ(k - 1) : 0.0);
    const int nseed = NCARRY + 32 + 2;
    for (int seed = 0; seed < nseed; seed++) {
      const int ps = seed - NCARRY;   // point seed: 0-7 Pu, 8-15 P0, 16-23 Pd, 24-31 Pn
      if (!has_pn && (ps >= 24 && ps < 32)) continue;
      if (!has_pn && seed == nseed - 1) continue;
      DeloState<D> d;
      Point<D> PnD;
      seed_point(s.Pu, d.Pu, ps);
      seed_point(s.P0, d.P0, ps - 8);
      seed_point(s.Pd, d.Pd, ps - 16);
      seed_point(Pn, PnD, ps - 24);
#pragma unroll
      for (int c = 0; c < NCARRY; c++) carry_ref(d, c) = D(carry_ref(s, c), c == seed ? R(1) : R(0));
      const D dzD(dzk, seed == nseed - 2 ? R(1) : R(0));
      const D dz1D(dzk1, seed == nseed - 1 ? R(1) : R(0));
      delo_step<D>(d, PnD, has_pn, dzD, dz1D, imuD);
      R r[4];
#pragma unroll
      for (int o = 0; o < 4; o++) {
        R acc = R(0);
#pragma unroll
        for (int c = 0; c < NCARRY; c++) acc += sbar[o][c] * carry_ref(d, c).d;
        r[o] = acc;
      }
      if (seed < NCARRY) {
#pragma unroll
        for (int o = 0; o < 4; o++) sbar_new[o][seed][threadIdx.x] = r[o];
      } else if (ps < 32) {
        const int kk = k + 1 - ps / 8, m = ps % 8;
#pragma unroll
        for (int o = 0; o < 4; o++) AP(o, kk, m) += r[o];
      } else {
        const int kk = seed == nseed - 2 ? k : k - 1;
#pragma unroll
        for (int o = 0; o < 4; o++) AZ(o, kk) += r[o];
      }
    }
#pragma unroll
    for (int o = 0; o < 4; o++)
#pragma unroll
      for (int c = 0; c < NCARRY; c++) sbar[o][c] = sbar_new[o][c][threadIdx.x];
  }

  // --- prologue: inputs are the points N-1 (Pu), N-2 (P0), N-3 (Pd),
  // |dz(N-1)|, |dz(N-2)| and the Planck values B(N-2), B(N-1)
  {
    Point<R> Pu, P0, Pd;
    load_point(pts, t, Nmax, N - 1, ncol, ic, Pu);
    load_point(pts, t, Nmax, N - 2, ncol, ic, P0);
    load_point(pts, t, Nmax, N - 3, ncol, ic, Pd);
    const R B2a = (R) planck(Tc[(int64_t) (N - 2) * ncol], lc[RHG_LC_HC_KLA], lc[RHG_LC_TWOHNU3_C2]);
    const R B2b = (R) planck(Tc[(int64_t) (N - 1) * ncol], lc[RHG_LC_HC_KLA], lc[RHG_LC_TWOHNU3_C2]);
    const R dz1 = (R) DZ(N - 1), dz2 = (R) DZ(N - 2);
    for (int seed = 0; seed < 28; seed++) {
      DeloState<D> d;
      Point<D> PdD;
      seed_point(Pu, d.Pu, seed);
      seed_point(P0, d.P0, seed - 8);
      seed_point(Pd, PdD, seed - 16);
      delo_prologue<D>(d, PdD, D(dz1, seed == 24 ? R(1) : R(0)), D(dz2, seed == 25 ? R(1) : R(0)),
                       imuD, D(B2a, seed == 26 ? R(1) : R(0)), D(B2b, seed == 27 ? R(1) : R(0)));
      R r[4];
#pragma unroll
      for (int o = 0; o < 4; o++) {
        R acc = R(0);
#pragma unroll
        for (int c = 0; c < NCARRY; c++) acc += sbar[o][c] * carry_ref(d, c).d;
        r[o] = acc;
      }
      if (seed < 24) {
        const int kk = N - 1 - seed / 8, m = seed % 8;
#pragma unroll
        for (int o = 0; o < 4; o++) AP(o, kk, m) += r[o];
      } else if (seed < 26) {
        const int kk = seed == 24 ? N - 1 : N - 2;
#pragma unroll
        for (int o = 0; o < 4; o++) AZ(o, kk) += r[o];
      } else {
#pragma unroll
        for (int o = 0; o < 4; o++) AB(o, seed - 26) = r[o];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 4. chain to the atmosphere parameters

// rf[((((p * Nmax + k) * 4 + o) * nout + io) * ncol + column]; chiref and
// dchiref are [column][depth] (as k_geometry writes chi_ref)
template <typename R>
__global__ void __launch_bounds__(64)
k_rf_chain(const DevModel *__restrict__ Mp, const int32_t *__restrict__ ndep,
           const int32_t *__restrict__ moving_arr, int ncol, int Nmax,
           const double *__restrict__ state, const double *__restrict__ dstate,
           const double *__restrict__ chiref, const double *__restrict__ dchiref,
           const R *__restrict__ cchi, const R *__restrict__ ceta,
           const R *__restrict__ dcchi, const R *__restrict__ dceta,
           const int32_t *__restrict__ out_ls, const int32_t *__restrict__ io_list, int nlist,
           int nout, double mu_d, int atm_scale, const R *__restrict__ aP,
           const R *__restrict__ aZ, const R *__restrict__ aB, R *__restrict__ rf) {
  RF_SETUP();
  using D = Dual<R>;
  const int has_line = M.lam_line_ptr[ls + 1] > M.lam_line_ptr[ls];
  const D muD = D((R) mu_d);
  const int64_t npts = (int64_t) Nmax * ncol;
  const int64_t tan_stride = (int64_t) state_nfield(M.ncont, M.nline) * npts;
  for (int k = 0; k < N; k++) {
    const double *stk = st0 + (int64_t) k * ncol;
    const int64_t ck = ((int64_t) io * Nmax + k) * ncol + ic;
    R ap[4][8];
#pragma unroll
    for (int o = 0; o < 4; o++)
#pragma unroll
      for (int m = 0; m < 8; m++) ap[o][m] = AP(o, k, m);
    for (int p = 0; p < RHG_RF_NPARAM; p++) {
      const double *dst = dstate + p * tan_stride + (int64_t) k * ncol + ic;
      auto F = [stk, dst, fs](int f) { return D((R) stk[f * fs], (R) dst[f * fs]); };
      const bool pT = p == RHG_RF_T;
      Point<D> P;
      form_point_f<D>(M, ls, has_line, F, muD, moving, D(cchi[ck], pT ? dcchi[ck] : R(0)),
                      D(ceta[ck], pT ? dceta[ck] : R(0)), P);
      R res[4];
#pragma unroll
      for (int o = 0; o < 4; o++) {
        R acc = R(0);
#pragma unroll
        for (int m = 0; m < 8; m++) acc += ap[o][m] * point_ref(P, m).d;
        res[o] = acc;
      }
      if (pT) {
        // path lengths on the tau500 scale: |dz(j)| = 2 (tau_j - tau_{j-1}) /
        // (chi_ref(j-1) + chi_ref(j)), j = k and k + 1
        if (atm_scale == 0) {
          const double dchi = dchiref[(int64_t) ic * Nmax + k];
#pragma unroll
          for (int jj = 0; jj < 2; jj++) {
            const int j = k + jj;
            if (j >= 1 && j <= N - 1) {
              const double den = chiref[(int64_t) ic * Nmax + j - 1] + chiref[(int64_t) ic * Nmax + j];
              const R dzdT = (R) (-DZ(j) / den * dchi);
#pragma unroll
              for (int o = 0; o < 4; o++) res[o] += AZ(o, j) * dzdT;
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
          const double dTdp = dst[RHG_CF_T * fs];
#pragma unroll
          for (int o = 0; o < 4; o++) res[o] += AB(o, k - (N - 2)) * (R) (dB * dTdp);
        }
      }
#pragma unroll
      for (int o = 0; o < 4; o++)
        rf[((((int64_t) p * Nmax + k) * 4 + o) * nout + io) * ncol + ic] = res[o];
    }
  }
}

#undef RF_SETUP
#undef DZ
#undef PTS
#undef CAR
#undef AP
#undef AZ
#undef AB

#define RHG_RF_INST(R)                                                                          \
  template __global__ void k_rf_forward<R>(const DevModel *, const int32_t *, const int32_t *,  \
                                           int, int, const double *, const int32_t *,          \
                                           const int32_t *, int, int, const R *, const R *,    \
                                           double, double *, R *, R *);                        \
  template __global__ void k_rf_reverse<R>(const DevModel *, const int32_t *, const int32_t *,  \
                                           int, int, const double *, const int32_t *,          \
                                           const int32_t *, int, double, const R *, const R *, \
                                           R *, R *, R *);                                     \
  template __global__ void k_rf_chain<R>(const DevModel *, const int32_t *, const int32_t *,    \
                                         int, int, const double *, const double *,             \
                                         const double *, const double *, const R *, const R *, \
                                         const R *, const R *, const int32_t *, const int32_t *, \
                                         int, int, double, int, const R *, const R *,          \
                                         const R *, R *);
RHG_RF_INST(double)
RHG_RF_INST(float)
#undef RHG_RF_INST

}  // namespace rhg
