// Stage 4 kernels in FP64 (the plain form), the reference's analytic
// d I / d log gf, and the work census.  The kernel templates and the FP32
// form's instantiations are in rhg_formal.hpp and rhg_formal32.hip.
#include "rhg_formal.hpp"

namespace rhg {

// d I / d log gf of the emergent intensity: the analytic response of
// rh/rhf1d/bezier_1D.c:416-536 (Piecewise_Bezier3_1D with get_atomic_rfs),
// for the wavelengths of io_list that the scalar Bezier solves (a moving
// column at a BEZIER wavelength; the others keep the zero that
// rh/rhf1d/formal.c:59-65 sets).  Parameter p perturbs the log gf of model
// line par_line[p] (< 0: no line, zero response).  The recursion is the
// reference's, not the exact derivative: Z = (-dchi I + deta) / chi with
// dchi = ln10 chi_l phi and deta = ln10 eta_l phi of that line's emergent-ray
// profile (rh/kurucz.c:696-699), carried with the intensity's Bezier weights
// and control-point clamps; the downwind I in Z is still the inward ray's
// (rh/rhf1d/formal.c:168-172 solves to_obs = 0 first into the same array),
// and in the last, linear interval Z enters with the opposite sign.  pyrh
// keeps one background array per wavelength (rh/readj.c:241-268), last
// written by the emergent ray's pass (rh/background.c:470-580), so the
// inward ray runs on the emergent ray's opacity and source function.  The
// inward ray (incident intensity at the top: 0, or ATMOS_ITOP's) is solved first into
// Iin[(t * Nmax + k) * ncol + ic]; each parameter then takes one emergent
// sweep.  Output rf[(ic * nout + io) * npar + p]; FP64 only.
__global__ void __launch_bounds__(64)
k_rf_loggf(const DevModel *__restrict__ Mp, const int32_t *__restrict__ ndep,
           const int32_t *__restrict__ moving_arr, int ncol, int Nmax,
           const double *__restrict__ state, const int32_t *__restrict__ out_ls,
           const int32_t *__restrict__ io_list, int nlist, int nout,
           const double *__restrict__ cchi, const double *__restrict__ ceta, double mu_d,
           int npar, const int32_t *__restrict__ par_line, double *__restrict__ Iin,
           double *__restrict__ rf) {
  using R = double;
  RAY_TILE();
  RAY_SETUP_BASE();
  if (column_solver(ldc(M.solver)[ls], moving) != RHG_SOLVER_BEZIER) return;
  const double zmu = 1.0 / mu_d;
  const int lb = line_field_base(M.ncont);
  double *Ii = Iin + (int64_t) t * Nmax * ncol + ic;
#define IIN(kk) Ii[(int64_t) (kk) * ncol]
#define POINT_IN(kk, P) POINT(kk, P)

  // inward ray, to_obs = 0: depth 0 down to N-1
  {
    Point<R> Pm, P0, Pp, Pn;
    POINT_IN(0, Pm);
    POINT_IN(1, P0);
    POINT_IN(2, Pp);
    // ZERO or IRRADIATED top boundary (rh/rhf1d/bezier_1D.c:370-380)
    const double Itop = lc[RHG_LC_ITOP];
    IIN(0) = Itop;
    double dsup = DZ(1) * zmu, dsdn = DZ(2) * zmu;
    double dchi_up = (P0.chi - Pm.chi) / dsup;
    double dchi_c = cent_deriv<R>(dsup, dsdn, Pm.chi, P0.chi, Pp.chi);
    double c1 = rmax0(P0.chi - (dsup / 3.0) * dchi_c);
    double c2 = rmax0(Pm.chi + (dsup / 3.0) * dchi_up);
    double dtau_uw = dsup * (P0.chi + Pm.chi + c1 + c2) * 0.25;
    double dS_up = (P0.S[0] - Pm.S[0]) / dtau_uw;
    double I_upw = Itop;
    for (int k = 1; k < N - 1; k++) {
      dsdn = DZ(k + 1) * zmu;
      double dchi_dn;
      if (N - 1 - k > 1) {
        POINT_IN(k + 2, Pn);
        const double dsdn2 = DZ(k + 2) * zmu;
        dchi_dn = cent_deriv<R>(dsdn, dsdn2, P0.chi, Pp.chi, Pn.chi);
      } else {
        dchi_dn = (Pp.chi - P0.chi) / dsdn;
      }
      c1 = rmax0(P0.chi + (dsdn / 3.0) * dchi_c);
      c2 = rmax0(Pp.chi - (dsdn / 3.0) * dchi_dn);
      const double dtau_dw = dsdn * (P0.chi + Pp.chi + c1 + c2) * 0.25;
      const double dt03 = dtau_uw / 3.0;
      const Bezier3<R> bz = bezier3(dtau_uw);
      const double dS_c = cent_deriv<R>(dtau_uw, dtau_dw, Pm.S[0], P0.S[0], Pp.S[0]);
      c1 = rmax0(P0.S[0] - dt03 * dS_c);
      c2 = rmax0(Pm.S[0] + dt03 * dS_up);
      const double I = I_upw * bz.eps + bz.alpha * P0.S[0] + bz.beta * Pm.S[0] + bz.gamma * c1 +
                       bz.theta * c2;
      IIN(k) = I;
      I_upw = I;
      dsup = dsdn;
      dchi_up = dchi_c;
      dchi_c = dchi_dn;
      dtau_uw = dtau_dw;
      dS_up = dS_c;
      Pm = P0;
      P0 = Pp;
      Pp = Pn;
    }
    double w[3];
    const double dtl = 0.5 * zmu * (P0.chi + Pm.chi) * DZ(N - 1);
    const double dS_uw = -(P0.S[0] - Pm.S[0]) / dtl;
    w3(dtl, w);
    IIN(N - 1) = (1.0 - w[0]) * I_upw + w[0] * P0.S[0] + w[1] * dS_uw;
  }

  const int p0 = ldc(M.lam_line_ptr)[ls], p1 = ldc(M.lam_line_ptr)[ls + 1];
  for (int p = 0; p < npar; p++) {
    const int lp = par_line[p];
    int ep = -1;
    for (int e = p0; e < p1; e++)
      if (ldc(M.lam_line_idx)[e] == lp) ep = e;
    if (lp < 0 || ep < 0) continue;
    // d chi and d eta of the parameter's line at depth kk
    auto dline = [&](int kk, double &a, double &b) {
      const double *sp = st0 + (int64_t) kk * ncol;
      auto F = [sp, fs](int f) { return sp[f * fs]; };
      double pq, pu, pv;
      const double phi = line_profile_f<double>(M, ep, lp, F, mu, F(RHG_CF_VLOS), moving, pq, pu, pv);
      a = b = 0.0;
      if (phi != 0.0) {
        a = F(lb + 4 * lp) * phi * LG10;
        b = F(lb + 4 * lp + 1) * phi * LG10;
      }
    };
    Point<R> Pu, P0, Pd, Pn;
    double au = 0, bu = 0, a0 = 0, b0 = 0, ad = 0, bd = 0, an = 0, bn = 0;
    double I_upw = 0, Iu = 0, dI_upw = 0, dZup = 0;
    double dtau_uw = 0, dsup = 0, dchi_up = 0, dchi_c = 0, dS_up = 0;
    for (int j = N - 1; j >= -1; j--) {
      if (j >= 0) {
        POINT(j, Pn);
        dline(j, an, bn);
      }
      if (j == N - 1) {
        Pu = Pn; au = an; bu = bn;
        continue;
      }
      if (j == N - 2) {
        P0 = Pn; a0 = an; b0 = bn;
        continue;
      }
      if (j == N - 3) {
        Pd = Pn; ad = an; bd = bn;
        const int k = N - 2;
        const double dtau0 = 0.5 * zmu * (Pu.chi + P0.chi) * DZ(N - 1);
        I_upw = B2[1] - (B2[0] - B2[1]) / dtau0;
        Iu = I_upw;
        dsup = DZ(k + 1) * zmu;
        const double dsdn = DZ(k) * zmu;
        dchi_up = (P0.chi - Pu.chi) / dsup;
        dchi_c = cent_deriv<R>(dsup, dsdn, Pu.chi, P0.chi, Pd.chi);
        const double c1 = rmax0(P0.chi - (dsup / 3.0) * dchi_c);
        const double c2 = rmax0(Pu.chi + (dsup / 3.0) * dchi_up);
        dtau_uw = dsup * (P0.chi + Pu.chi + c1 + c2) * 0.25;
        dS_up = (P0.S[0] - Pu.S[0]) / dtau_uw;
        dI_upw = 0.0;
        const double Zk = -a0 / P0.chi * IIN(k) + b0 / P0.chi;
        const double Zkm1 = -au / Pu.chi * Iu + bu / Pu.chi;
        dZup = (Zk - Zkm1) / dtau_uw;
        continue;
      }
      const int k = j + 2;
      const double dsdn = DZ(k) * zmu;
      double dchi_dn;
      if (j >= 0) {
        const double dsdn2 = DZ(k - 1) * zmu;
        dchi_dn = cent_deriv<R>(dsdn, dsdn2, P0.chi, Pd.chi, Pn.chi);
      } else {
        dchi_dn = (Pd.chi - P0.chi) / dsdn;
      }
      double c1 = rmax0(P0.chi + (dsdn / 3.0) * dchi_c);
      double c2 = rmax0(Pd.chi - (dsdn / 3.0) * dchi_dn);
      const double dtau_dw = dsdn * (P0.chi + Pd.chi + c1 + c2) * 0.25;
      const double dt03 = dtau_uw / 3.0;
      const Bezier3<R> bz = bezier3(dtau_uw);
      const double dS_c = cent_deriv<R>(dtau_uw, dtau_dw, Pu.S[0], P0.S[0], Pd.S[0]);
      c1 = rmax0(P0.S[0] - dt03 * dS_c);
      c2 = rmax0(Pu.S[0] + dt03 * dS_up);
      const double I = I_upw * bz.eps + bz.alpha * P0.S[0] + bz.beta * Pu.S[0] + bz.gamma * c1 +
                       bz.theta * c2;
      // the response at k, with the inward ray's intensity at k - 1
      const double Zk = (-a0 * I + b0) / P0.chi;
      const double Zkm1 = (-au * Iu + bu) / Pu.chi;
      const double Zkp1 = (-ad * IIN(k - 1) + bd) / Pd.chi;
      const double dZk = cent_deriv<R>(dtau_uw, dtau_dw, Zkm1, Zk, Zkp1);
      c1 = rmax0(Zk - dt03 * dZk);
      c2 = rmax0(Zkm1 + dt03 * dZup);
      const double dI = dI_upw * bz.eps + bz.alpha * Zk + bz.beta * Zkm1 + bz.gamma * c1 +
                        bz.theta * c2;
      I_upw = I;
      Iu = I;
      dI_upw = dI;
      dZup = dZk;
      Pu = P0; au = a0; bu = b0;
      P0 = Pd; a0 = ad; b0 = bd;
      if (j >= 0) { Pd = Pn; ad = an; bd = bn; }
      dsup = dsdn;
      dchi_up = dchi_c;
      dchi_c = dchi_dn;
      dtau_uw = dtau_dw;
      dS_up = dS_c;
    }
    double w[3];
    const double dtl = 0.5 * zmu * (P0.chi + Pu.chi) * DZ(1);
    const double dS_uw = -(P0.S[0] - Pu.S[0]) / dtl;
    w3(dtl, w);
    const double I0 = (1.0 - w[0]) * I_upw + w[0] * P0.S[0] + w[1] * dS_uw;
    const double Zk = a0 / P0.chi * I0 - b0 / P0.chi;
    const double Zkm1 = au / Pu.chi * Iu - bu / Pu.chi;
    const double dZk = -(Zk - Zkm1) / dtl;
    rf[((int64_t) ic * nout + io) * npar + p] = (1.0 - w[0]) * dI_upw + w[0] * Zk + w[1] * dZk;
  }
#undef IIN
#undef POINT_IN
}

#undef POINT
#undef Z
#undef DZ
#undef RAY_SETUP
#undef RAY_SETUP_BASE
#undef RAY_TILE
#undef RAY_TILE_GRID

#define RHG_KFORMAL(KNAME, R)                                                                 \
  template __global__ void KNAME<R>(const DevModel *, const int32_t *, const int32_t *, int, int, \
                                    const double *, const int32_t *, const int32_t *, int, int,  \
                                    const R *, const R *, double, double *, const int32_t *, \
                                    const int32_t *, const int32_t *);
#define RHG_KSTOKES(R, MO, LT)                                                                \
  template __global__ void k_formal_stokes<R, MO, LT>(const DevModel *, const int32_t *,            \
                                                      const int32_t *, int, int, const double *,    \
                                                      const int32_t *, const int32_t *, int, int,   \
                                                      const R *, const R *, double, double *,       \
                                                      const int32_t *, const int32_t *,             \
                                                      const int32_t *, int, int);
RHG_KSTOKES(double, false, false)
RHG_KSTOKES(double, true, false)
RHG_KSTOKES(double, false, true)
RHG_KSTOKES(double, true, true)
#undef RHG_KSTOKES
RHG_KFORMAL(k_formal_scalar, double)
#undef RHG_KFORMAL

// Work census: one thread per (column, output wavelength), the mapping of
// k_formal; counts per ray in census[(ic * nout + io) * RHG_NWORK + slot].
__global__ void __launch_bounds__(64)
k_census(const DevModel *__restrict__ Mp, const int32_t *__restrict__ ndep, const int32_t *__restrict__ moving_arr,
         int ncol, int Nmax, const double *__restrict__ state, int nfield,
         const int32_t *__restrict__ out_ls, int nout, double mu, int32_t *__restrict__ census) {
  const DevModel &M = *Mp;
  const int ic = blockIdx.x * blockDim.x + threadIdx.x;
  const int io = blockIdx.y;
  if (ic >= ncol || io >= nout) return;
  const int N = ndep[ic], ls = out_ls[io], moving = moving_arr[ic];
  const int64_t fs = (int64_t) Nmax * ncol;
  const double *st0 = state + ic;
  int32_t cnt[RHG_NWORK];
  for (int i = 0; i < RHG_NWORK; i++) cnt[i] = 0;
  const int solver = column_solver(ldc(M.solver)[ls], moving);
  const int slot = solver == RHG_SOLVER_STOKES ? RHG_WORK_STOKES
                 : solver == RHG_SOLVER_BEZIER ? RHG_WORK_BEZIER : RHG_WORK_FEAUTRIER;
  const int has_line = ldc(M.lam_line_ptr)[ls + 1] > ldc(M.lam_line_ptr)[ls];
  for (int k = 1; k < N; k++) {
    cnt[slot]++;
    if (has_line) line_census(M, ls, st0 + (int64_t) k * ncol, fs, mu, moving, cnt);
  }
  int32_t *o = census + ((int64_t) ic * nout + io) * RHG_NWORK;
  for (int i = 0; i < RHG_NWORK; i++) o[i] = cnt[i];
}

// Per-column sums of the census: thread per column.
__global__ void k_census_sum(const int32_t *__restrict__ census, int ncol, int nout,
                             int64_t *__restrict__ col_sum) {
  const int ic = blockIdx.x * blockDim.x + threadIdx.x;
  if (ic >= ncol) return;
  int64_t acc[RHG_NWORK];
  for (int i = 0; i < RHG_NWORK; i++) acc[i] = 0;
  for (int io = 0; io < nout; io++)
    for (int i = 0; i < RHG_NWORK; i++) acc[i] += census[((int64_t) ic * nout + io) * RHG_NWORK + i];
  for (int i = 0; i < RHG_NWORK; i++) col_sum[(int64_t) ic * RHG_NWORK + i] = acc[i];
}

}  // namespace rhg
