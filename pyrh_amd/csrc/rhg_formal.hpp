// Stage 4: emergent-ray formal solution.  One thread per (column,
// wavelength); a wavefront is 64 columns at one wavelength, so the
// wavelength's line list, Zeeman pattern and solver are wave-uniform and the
// Humlicek branch taken by a lane follows its own column's Doppler offset
// (columns at one wavelength mostly share a branch, where wavelengths
// across a line do not).  Opacity and source functions are formed depth by
// depth, two points ahead of the recursion, from the stage-3 continuum and
// the per-depth line constants of the column state; nothing per depth is
// kept except the rolling window the recursion needs.
//
// Precision R: double is the plain form; float is the FP32 form.  In the
// FP32 form the continuum is stored in float and the Voigt/Faraday-Voigt
// functions of every Zeeman component (the bulk of the work) are evaluated
// in float.  The polarized kernel sums and combines those profiles, forms
// the source functions and propagation matrix and runs the DELO recursion in
// FP64: at the core of a strong Zeeman line |q| -> 1 and the weak
// polarization mode's opacity chi_I (1 - |q|) and source (S_I - q.S_QUV)
// are differences of nearly equal terms, so a float q or S puts 1e-4 I_c
// errors into Q, U, V.  The unpolarized kernel keeps float points; its
// Bezier and w3 weights, path lengths and the Feautrier recursion are FP64.
// The FP32 instantiations are compiled with FMA contraction, the FP64 ones
// without (they follow the reference's operation order).
#pragma once

#include "rhg_delo.hpp"

namespace rhg {

// The solver of a wavelength in one column: BEZIER is angle-independent in a
// static column and takes Feautrier (rh/rhf1d/formal.c:98-101, 287-310);
// BEZIER_ANGLE (a polarizable background line outside FULL_STOKES) stays
// with the scalar Bezier.
__device__ __forceinline__ int column_solver(int solver, int moving) {
  if (solver == RHG_SOLVER_BEZIER_ANGLE) return RHG_SOLVER_BEZIER;
  return (solver == RHG_SOLVER_BEZIER && !moving) ? RHG_SOLVER_FEAUTRIER : solver;
}

// form_point over the column state at st (stride fs)
template <typename R>
__device__ __forceinline__ void form_point(const DevModel &M, int ls, int has_line,
                                           const double *st, int64_t fs, R mu, int moving,
                                           R cchi, R ceta, Point<R> &p) {
  auto F = [st, fs](int f) { return (R) st[f * fs]; };
  form_point_f<R>(M, ls, has_line, F, mu, moving, cchi, ceta, p);
}


// Workgroup -> (column block cb, output wavelength) of the formal kernels.
// Without wavelength grids (blk_grid == nullptr): entry t < nlist of
// io_list.  With grids, the columns of block cb all belong to grid g =
// blk_grid[cb] (-1: a block without columns), whose wavelengths are entries
// gptr[g] .. gptr[g + 1] - 1 of io_list, and nlist is the longest grid's
// count.  Returns -1 for a workgroup with nothing to solve.
__device__ __forceinline__ int ray_io(const int32_t *io_list, int nlist, const int32_t *blk_grid,
                                      const int32_t *gptr, int ncol, int &cb) {
  int t;
  xcd_tile(blockIdx.x, nlist, cb, t);
  if (!blk_grid) return io_list[t];
  if (cb * 64 >= ncol) return -1;
  const int g = ldc(blk_grid)[cb];
  if (g < 0) return -1;
  const int e = ldc(gptr)[g] + t;
  return e < ldc(gptr)[g + 1] ? ldc(io_list)[e] : -1;
}

// workgroup -> (column block, entry t of the wavelength list io_list)
#define RAY_TILE()                                                                       \
  int cb, t;                                                                             \
  xcd_tile(blockIdx.x, nlist, cb, t);                                                    \
  const int io = io_list[t]
// the same through ray_io (kernels that take wavelength grids)
#define RAY_TILE_GRID()                                                                  \
  int cb;                                                                                \
  const int io = ray_io(io_list, nlist, blk_grid, gptr, ncol, cb);                       \
  if (io < 0) return

// Common set-up of one ray (after RAY_TILE or RAY_TILE_GRID): lane -> column.
#define RAY_SETUP_BASE()                                                                 \
  const DevModel &M = *Mp;                                                               \
  const int ic = cb * 64 + threadIdx.x;                                                  \
  if (ic >= ncol) return;                                                                \
  const int N = ndep[ic];                                                                \
  const int ls = out_ls[io];                                                             \
  const int moving = moving_arr[ic];                                                     \
  const int64_t fs = (int64_t) Nmax * ncol;                                              \
  const double *st0 = state + ic;                                                        \
  const double *zc = st0 + (int64_t) RHG_CF_HEIGHT * fs;                                 \
  const double *dzc = st0 + (int64_t) RHG_CF_DZ * fs;                                    \
  const double *Tc = st0 + (int64_t) RHG_CF_T * fs;                                      \
  const R *cc = cchi + (int64_t) io * Nmax * ncol + ic;                                  \
  const R *ce = ceta + (int64_t) io * Nmax * ncol + ic;                                  \
  const int has_line = ldc(M.lam_line_ptr)[ls + 1] > ldc(M.lam_line_ptr)[ls];                      \
  const double *lc = M.lamconst + (int64_t) ls * RHG_LC_NCONST;                          \
  const R mu = (R) mu_d;                                                                 \
  double B2[2];                                                                          \
  B2[0] = planck(Tc[(int64_t) (N - 2) * ncol], lc[RHG_LC_HC_KLA], lc[RHG_LC_TWOHNU3_C2]); \
  B2[1] = planck(Tc[(int64_t) (N - 1) * ncol], lc[RHG_LC_HC_KLA], lc[RHG_LC_TWOHNU3_C2]);

// the same, with the column's Stokes output for wavelength io (slot ic holds
// the caller's column perm[ic], or ic without a permutation; a negative
// entry marks a slot that only pads a grid's columns to whole blocks, which
// writes nothing)
#define RAY_SETUP()                                                                      \
  RAY_SETUP_BASE();                                                                      \
  const int oc = perm ? perm[ic] : ic;                                                   \
  if (oc < 0) return;                                                                    \
  double *out = stokes + (int64_t) oc * 4 * nout + io;

#define Z(kk) zc[(int64_t) (kk) * ncol]
// |z[kk-1] - z[kk]| (kk >= 1), from the column state
#define DZ(kk) dzc[(int64_t) (kk) * ncol]
#define POINT(kk, P) form_point<R>(M, ls, has_line, st0 + (int64_t) (kk) * ncol, fs, mu, moving, \
                                   cc[(int64_t) (kk) * ncol], ce[(int64_t) (kk) * ncol], P)

// Polarized cubic DELO-Bezier, rh/rhf1d/bezier_1D.c:52-300 (to_obs: depth
// from N-1 to 0), for the wavelengths of io_list (solver STOKES).  The loop
// forms one new point per iteration (index j, the only place the opacity is
// evaluated); the first three iterations fill the window, each later one
// advances the recursion at depth k = j + 2, and j = -1 takes the step at
// k = 1 whose downwind derivative is one-sided.  MO: with the
// magneto-optical terms (rh/stokesopac.c:56-86, rh/kurucz.c:711-715).
// LT: the wavelength's line list (lines, Doppler offsets, Zeeman
// components) is staged in LDS once per wavefront, tab_maxl lines and
// tab_maxc components at most (the maxima over the model's wavelengths), so
// that the depth loop reads those constants from LDS rather than from
// global memory at every depth.
template <typename R, bool MO, bool LT>
__global__ void __launch_bounds__(64)
k_formal_stokes(const DevModel *__restrict__ Mp, const int32_t *__restrict__ ndep,
                const int32_t *__restrict__ moving_arr, int ncol, int Nmax,
                const double *__restrict__ state, const int32_t *__restrict__ out_ls,
                const int32_t *__restrict__ io_list, int nlist, int nout,
                const R *__restrict__ cchi, const R *__restrict__ ceta, double mu_d,
                double *__restrict__ stokes, const int32_t *__restrict__ perm,
                const int32_t *__restrict__ blk_grid, const int32_t *__restrict__ gptr,
                int tab_maxl, int tab_maxc) {
  RAY_TILE_GRID();
  // the line table first: every lane of the wavefront stages it
  auto tab = [&]() {
    const int ls0 = out_ls[io];
    if constexpr (LT) {
      extern __shared__ __align__(8) char lds_tab[];
      return stage_lines(*Mp, ls0, lds_tab, tab_maxl, tab_maxc);
    } else {
      const int p0 = Mp->lam_line_ptr[ls0];
      return LineTabG(*Mp, p0, Mp->lam_line_ptr[ls0 + 1] - p0);
    }
  }();
  RAY_SETUP();
  (void) mu;
  const double imu = 1.0 / mu_d;
  DeloState<double> st;
  Point<double> Pn;
  for (int j = N - 1; j >= -1; j--) {
    if (j >= 0)
      form_point_vt<R, R, MO>(tab, M.ncont, has_line, st0 + (int64_t) j * ncol, fs, mu_d, moving,
                              cc[(int64_t) j * ncol], ce[(int64_t) j * ncol], Pn);
    if (j == N - 1) {
      st.Pu = Pn;
    } else if (j == N - 2) {
      st.P0 = Pn;
    } else if (j == N - 3) {
      delo_prologue<double>(st, Pn, DZ(N - 1), DZ(N - 2), imu, B2[0], B2[1]);
    } else {
      const int k = j + 2;
      if constexpr (MO) delo_step_mo<double>(st, Pn, j >= 0, DZ(k), j >= 0 ? DZ(k - 1) : 0.0, imu);
      else delo_step<double>(st, Pn, j >= 0, DZ(k), j >= 0 ? DZ(k - 1) : 0.0, imu);
    }
  }
  double x[4];
  if constexpr (MO) delo_final_mo<double>(st, DZ(1), imu, x);
  else delo_final<double>(st, DZ(1), imu, x);
#pragma unroll
  for (int i = 0; i < 4; i++) out[(int64_t) i * nout] = x[i];
}

// Unpolarized rays (solver BEZIER or FEAUTRIER) for the wavelengths of
// io_list: the scalar cubic Bezier of rh/rhf1d/bezier_1D.c:306-541 for a
// moving column, else Feautrier, rh/rhf1d/feautrier.c:56-202 (STANDARD order,
// top boundary from ATMOS_ITOP or zero, thermalised bottom, FP64).  A BEZIER
// wavelength is solved with Feautrier in a static column, as
// rh/rhf1d/formal.c:98-101 decides (column_solver).  Q = U = V = 0.
template <typename R>
__global__ void __launch_bounds__(64)
k_formal_scalar(const DevModel *__restrict__ Mp, const int32_t *__restrict__ ndep,
                const int32_t *__restrict__ moving_arr, int ncol, int Nmax,
                const double *__restrict__ state, const int32_t *__restrict__ out_ls,
                const int32_t *__restrict__ io_list, int nlist, int nout,
                const R *__restrict__ cchi, const R *__restrict__ ceta, double mu_d,
                double *__restrict__ stokes, const int32_t *__restrict__ perm,
                const int32_t *__restrict__ blk_grid, const int32_t *__restrict__ gptr) {
  RAY_TILE_GRID();
  RAY_SETUP();
  const int solver = column_solver(ldc(M.solver)[ls], moving);
  Point<R> Pn;
  if (solver == RHG_SOLVER_BEZIER) {
    const double zmu = 1.0 / mu_d;
    Point<R> Pu, P0, Pd;
    R I_upw = R(0), dtau_uw = R(0), dsup = R(0), dchi_up = R(0), dchi_c = R(0), dS_up = R(0);
    for (int j = N - 1; j >= -1; j--) {
      if (j >= 0) POINT(j, Pn);
      if (j >= N - 3) {
        if (j == N - 1) {
          Pu = Pn;
        } else if (j == N - 2) {
          P0 = Pn;
        } else {
          Pd = Pn;
          const int k = N - 2;
          const double dtau0 = 0.5 * zmu * (double) (Pu.chi + P0.chi) * DZ(N - 1);
          I_upw = (R) (B2[1] - (B2[0] - B2[1]) / dtau0);
          dsup = (R) (DZ(k + 1) * zmu);
          const R dsdn = (R) (DZ(k) * zmu);
          dchi_up = rdiv(P0.chi - Pu.chi, dsup);
          dchi_c = cent_deriv<R>(dsup, dsdn, Pu.chi, P0.chi, Pd.chi);
          const R c1 = rmax0(P0.chi - (dsup * R(1.0 / 3.0)) * dchi_c);
          const R c2 = rmax0(Pu.chi + (dsup * R(1.0 / 3.0)) * dchi_up);
          dtau_uw = dsup * (P0.chi + Pu.chi + c1 + c2) * R(0.25);
          dS_up = rdiv(P0.S[0] - Pu.S[0], dtau_uw);
        }
        continue;
      }
      const int k = j + 2;
      const R dsdn = (R) (DZ(k) * zmu);
      R dchi_dn;
      if (j >= 0) {
        const R dsdn2 = (R) (DZ(k - 1) * zmu);
        dchi_dn = cent_deriv<R>(dsdn, dsdn2, P0.chi, Pd.chi, Pn.chi);
      } else {
        dchi_dn = rdiv(Pd.chi - P0.chi, dsdn);
      }
      R c1 = rmax0(P0.chi + (dsdn * R(1.0 / 3.0)) * dchi_c);
      R c2 = rmax0(Pd.chi - (dsdn * R(1.0 / 3.0)) * dchi_dn);
      const R dtau_dw = dsdn * (P0.chi + Pd.chi + c1 + c2) * R(0.25);
      const Bezier3<R> bz = bezier3(dtau_uw);
      const R dt03 = dtau_uw * R(1.0 / 3.0);
      const R dS_c = cent_deriv<R>(dtau_uw, dtau_dw, Pu.S[0], P0.S[0], Pd.S[0]);
      c1 = rmax0(P0.S[0] - dt03 * dS_c);
      c2 = rmax0(Pu.S[0] + dt03 * dS_up);
      I_upw = I_upw * bz.eps + bz.alpha * P0.S[0] + bz.beta * Pu.S[0] + bz.gamma * c1 +
              bz.theta * c2;
      Pu = P0;
      P0 = Pd;
      if (j >= 0) Pd = Pn;
      dsup = dsdn;
      dchi_up = dchi_c;
      dchi_c = dchi_dn;
      dtau_uw = dtau_dw;
      dS_up = dS_c;
    }
    double w[3];
    const double dtl = 0.5 * zmu * (double) (P0.chi + Pu.chi) * DZ(1);
    const double dS_uw = -(double) (P0.S[0] - Pu.S[0]) / dtl;
    w3(dtl, w);
    out[0] = (1.0 - w[0]) * (double) I_upw + w[0] * (double) P0.S[0] + w[1] * dS_uw;
  } else {
    // The back substitution for P[0] is accumulated in the forward sweep:
    // P[0] = sum_k ztmp[k] prod_{j<k} 1/(1+F[j]) + P[N-1] prod_{j<N-1} 1/(1+F[j]).
    const double zmu = 0.5 / mu_d;
    double chi_prev = 0.0, S_prev = 0.0, dtau_prev = 0.0, F = 0.0, ztmp = 0.0, P0acc = 0.0,
           prod = 1.0;
    for (int j = 0; j < N; j++) {
      POINT(j, Pn);
      const double chi_j = (double) Pn.chi, S_j = (double) Pn.S[0];
      if (j == 1) {
        dtau_prev = zmu * (chi_prev + chi_j) * (Z(0) - Z(1));
        const double abc0 = 1.0 + 2.0 * 1.0 / dtau_prev;
        const double C10 = 2.0 / (dtau_prev * dtau_prev);
        // h0: the intensity irradiating the top (0 without ATMOS_ITOP), feautrier.c:84-97
        const double Stmp = S_prev + 2.0 * lc[RHG_LC_ITOP] / ((1.0 + 0.0) * dtau_prev);
        F = abc0 / C10;
        ztmp = Stmp / (abc0 + C10);
        P0acc = ztmp;
        prod = 1.0 / (1.0 + F);
      } else if (j >= 2) {
        // row k = j - 1 with its downwind point j
        const int k = j - 1;
        const double dtau_k = zmu * (chi_prev + chi_j) * (Z(k) - Z(k + 1));
        const double dtau_mid = 0.5 * (dtau_k + dtau_prev);
        const double A1 = 1.0 / (dtau_mid * dtau_prev);
        const double C1 = 1.0 / (dtau_mid * dtau_k);
        const double Fn = (1.0 + A1 * F / (1.0 + F)) / C1;
        ztmp = (S_prev + A1 * ztmp) / (C1 * (1.0 + Fn));
        F = Fn;
        P0acc += ztmp * prod;
        prod *= 1.0 / (1.0 + F);
        dtau_prev = dtau_k;
      }
      chi_prev = chi_j;
      S_prev = S_j;
    }
    const double hN = B2[1] - (B2[0] - B2[1]) / dtau_prev;
    const double abcN = 1.0 + 2.0 * 1.0 / dtau_prev;
    const double A1N = 2.0 / (dtau_prev * dtau_prev);
    const double StmpN = S_prev + 2.0 * hN / ((1.0 + 0.0) * dtau_prev);
    const double PN = (StmpN + A1N * ztmp) / (abcN + A1N * (F / (1.0 + F)));
    const double P0 = P0acc + PN * prod;
    out[0] = (1.0 + 1.0) * P0 - lc[RHG_LC_ITOP] / (1.0 + 0.0);
  }
  out[nout] = 0.0;
  out[2 * (int64_t) nout] = 0.0;
  out[3 * (int64_t) nout] = 0.0;
}
}  // namespace rhg
