// gfx950 kernels for batched LTE Stokes synthesis: column state, geometry
// and continuum.  Every kernel maps consecutive lanes to consecutive
// columns, so loads and stores of the column state
// (state[(field * Nmax + k) * ncol + column]) are coalesced.
//
// Stage 1  k_state      one thread per (column, depth): unit conversion,
//                       B projection, Saha-Boltzmann populations of every
//                       model atom, chemical equilibrium (Newton), Kurucz
//                       element populations, per-line opacity constants
//                       and the background opacity at the reference
//                       wavelength.  Writes the column state.
// Stage 2  k_geometry   one thread per column: tau500 -> height
//                       (convertScales, a serial scan of Ndep).
// Stage 3  k_continuum  one thread per (column, depth), all output
//                       wavelengths: continuum opacity and emissivity, the
//                       bound-free sums register-tiled over BF_TILE
//                       wavelengths.
// Stage 4  k_formal     (rhg_formal.hip) one thread per (column, wavelength).
#include "rhg_kernels.hpp"
#include "rhg_equil.hpp"

namespace rhg {

// ---------------------------------------------------------------------------
// Stage 1: column state

// continuum (absorption) at one wavelength and depth; the sum order of
// rh/background.c:329-465.  Column-state pointer st points at depth k
// with stride `fs` between fields.
__device__ __forceinline__ void continuum_point(const DevModel &M, int ls, const double *st,
                                                int64_t fs, double &chi_ai, double &eta_ai) {
  const double *lc = M.lamconst + (int64_t) ls * RHG_LC_NCONST;
  const double T = st[RHG_CF_T * fs], ne = st[RHG_CF_NE * fs];
  const double C0 = Q_ELECTRON * Q_ELECTRON / (4.0 * PI * EPSILON_0) / sqrt(M_ELECTRON);
  const double sigma_ff = 4.0 / 3.0 * sqrt(2.0 * PI / (3.0 * KBOLTZMANN)) * (C0 * C0 * C0) /
                          (HPLANCK * CLIGHT);
  double chi, eta, stim, Bnu;
  chi_ai = 0.0;
  eta_ai = 0.0;
  Bnu = planck(T, lc[RHG_LC_HC_KLA], lc[RHG_LC_TWOHNU3_C2]);
  const double nHmin = st[RHG_CF_NHMIN * fs], nH0 = st[RHG_CF_NH0 * fs], np = st[RHG_CF_NP * fs];
  if (lc[RHG_LC_HMIN_BF_ON] != 0.0) {
    stim = exp(-lc[RHG_LC_HC_KLA] / T);
    chi_ai += nHmin * (1.0 - stim) * lc[RHG_LC_HMIN_BF_ALPHA];
    eta_ai += nHmin * lc[RHG_LC_TWOHNU3_C2] * stim * lc[RHG_LC_HMIN_BF_ALPHA];
  }
  if (lc[RHG_LC_HMIN_FF_LIDX] >= 0.0) {
    double ti = clamp_index(NTHETA, M.hminff_theta, THETA0 / T);
    double pe = ne * KBOLTZMANN * T;
    double kappa = bilinear(NTHETA, NFF, M.hminff_kappa, ti, lc[RHG_LC_HMIN_FF_LIDX]);
    chi = (nH0 * 1.0E-29) * pe * kappa;
  } else {
    double Ck = (KBOLTZMANN * THETA0 * 1.0E-32), theta_n = 1.0, sqrt_theta = sqrt(THETA0 / T);
    chi = 0.0;
    for (int n = 1; n < 6; n++) {
      theta_n *= sqrt_theta;
      chi += theta_n * lc[RHG_LC_HFF_LONG0 + n - 1];
    }
    chi *= nH0 * (ne * Ck);
  }
  chi_ai += chi;
  eta_ai += chi * Bnu;
  // the H- fudge factor scales what has been summed so far, rh/background.c:362-369
  chi_ai *= lc[RHG_LC_HMIN_FUDGE];
  eta_ai *= lc[RHG_LC_HMIN_FUDGE];
  if (lc[RHG_LC_OH_EIDX] >= 0.0) {
    if (T < ldc(M.oh_T)[0] || T > ldc(M.oh_T)[NTOH - 1]) { chi = 0.0; eta = 0.0; }
    else {
      int i2 = locate_asc(NTOH, M.oh_T, T);
      double ti = (i2 == NTOH - 1) ? (double) i2
                                   : (double) i2 + (T - ldc(M.oh_T)[i2]) / (ldc(M.oh_T)[i2 + 1] - ldc(M.oh_T)[i2]);
      double kappa = exp(LG10 * bilinear(NTOH, NEOH, M.oh_cross, ti, lc[RHG_LC_OH_EIDX])) *
                     (CM_TO_M * CM_TO_M);
      stim = exp(-lc[RHG_LC_HC_KLA] / T);
      double nOH = st[RHG_CF_NOH * fs];
      chi = nOH * (1.0 - stim) * kappa;
      eta = nOH * lc[RHG_LC_TWOHNU3_C2] * stim * kappa;
    }
    chi_ai += chi;
    eta_ai += eta;
  }
  if (lc[RHG_LC_CH_EIDX] >= 0.0) {
    if (T < ldc(M.ch_T)[0] || T > ldc(M.ch_T)[NTCH - 1]) { chi = 0.0; eta = 0.0; }
    else {
      int i2 = locate_asc(NTCH, M.ch_T, T);
      double ti = (i2 == NTCH - 1) ? (double) i2
                                   : (double) i2 + (T - ldc(M.ch_T)[i2]) / (ldc(M.ch_T)[i2 + 1] - ldc(M.ch_T)[i2]);
      double kappa = exp(LG10 * bilinear(NTCH, NECH, M.ch_cross, ti, lc[RHG_LC_CH_EIDX])) *
                     (CM_TO_M * CM_TO_M);
      stim = exp(-lc[RHG_LC_HC_KLA] / T);
      double nCH = st[RHG_CF_NCH * fs];
      chi = nCH * (1.0 - stim) * kappa;
      eta = nCH * lc[RHG_LC_TWOHNU3_C2] * stim * kappa;
    }
    chi_ai += chi;
    eta_ai += eta;
  }
  // bound-free: H continua first, then metals
  const int cp0 = ldc(M.lam_cont_ptr)[ls], cp1 = ldc(M.lam_cont_ptr)[ls + 1];
  double expla = exp(-lc[RHG_LC_BF_HC_KLA] / T);
  {
    bool any = false;
    chi = eta = 0.0;
    for (int e = cp0; e < cp1; e++) {
      int ci = ldc(M.lam_cont_idx)[e];
      if (ldc(M.cont_atom)[ci] != 0) continue;
      any = true;
      double sigma = ldc(M.lam_cont_alpha)[e];
      double ni = st[cont_field(ci) * fs];
      chi += sigma * (1.0 - expla) * ni;
      eta += lc[RHG_LC_BF_TWOHNU3_C2] * expla * sigma * ni;
    }
    if (any) { chi_ai += chi; eta_ai += eta; }
  }
  {
    double y = (lc[RHG_LC_GFF_CY] * T) / (HPLANCK * CLIGHT), x3 = lc[RHG_LC_GFF_X3];
    double gIII = 1.0 + 0.1728 * x3 * (1.0 + y) - 0.0496 * (x3 * x3) * (1.0 + (1.0 + y) * 0.33333333 * y);
    double g_ff = (gIII > 1.0) ? gIII : 1.0;
    stim = 1.0 - exp(-lc[RHG_LC_HC_KLA] / T);
    chi = sigma_ff / sqrt(T) * lc[RHG_LC_HFF_NU3] * ne * np * stim * g_ff;
    chi_ai += chi;
    eta_ai += chi * Bnu;
  }
  if (lc[RHG_LC_H2P_LIDX] >= 0.0) {
    double ti = clamp_index(NTEMP_H2P, M.h2pff_temp, T);
    double kappa = bilinear(NTEMP_H2P, NFF_H2P, M.h2pff_kappa, ti, lc[RHG_LC_H2P_LIDX]);
    chi = (nH0 * 1.0E-29) * (np * 1.0E-20) * kappa;
    chi_ai += chi;
    eta_ai += chi * Bnu;
  }
  if (lc[RHG_LC_H2M_LIDX] >= 0.0) {
    double nH2 = st[RHG_CF_NH2 * fs];
    if (nH2 > 0.0) {
      double ti = clamp_index(NTHETA_H2, M.h2mff_theta, THETA0 / T);
      double pe = ne * KBOLTZMANN * T;
      double kappa = bilinear(NTHETA_H2, NFF_H2, M.h2mff_kappa, ti, lc[RHG_LC_H2M_LIDX]);
      chi = (nH2 * 1.0E-29) * pe * kappa;
    } else chi = 0.0;
    chi_ai += chi;
    eta_ai += chi * Bnu;
  }
  chi = eta = 0.0;
  for (int e = cp0; e < cp1; e++) {
    int ci = ldc(M.lam_cont_idx)[e];
    if (ldc(M.cont_atom)[ci] == 0) continue;
    double alpha_la = ldc(M.lam_cont_alpha)[e];
    double ni = st[cont_field(ci) * fs];
    chi += alpha_la * (1.0 - expla) * ni;
    eta += lc[RHG_LC_BF_TWOHNU3_C2] * expla * alpha_la * ni;
  }
  chi_ai += chi * 1.0;
  eta_ai += eta * 1.0;
}

// Three waves per SIMD: 168 VGPRs (4 spilled) and 2 Nmol doubles of LDS
// per thread.
template <int NN>
__global__ void __launch_bounds__(64, 3)
k_state(const DevModel *__restrict__ Mp, const double *__restrict__ atmos, const int32_t *__restrict__ ndep,
        const int32_t *__restrict__ perm, int ncol, int Nmax, double mu, int atm_scale,
        double *__restrict__ state, int nfield) {
  const DevModel &M = *Mp;
  // workgroup -> (column block, depth): the depths of a column block run one
  // after another on one XCD, so the atmosphere's cache lines (16 depths of
  // one row of one column) are fetched into that XCD's L2 once
  int cb, k;
  xcd_tile(blockIdx.x, Nmax, cb, k);
  const int ic = cb * 64 + (int) threadIdx.x;
  if (ic >= ncol) return;
  if (k >= ndep[ic]) return;
  // slot ic of the batch holds the caller's column perm[ic] (identity without
  // perm); a slot that pads a grid's columns to whole blocks holds -1 - c, a
  // copy of column c whose results are not written
  int src = perm ? perm[ic] : ic;
  if (src < 0) src = -1 - src;
  const double *atm = atmos + (int64_t) src * 9 * Nmax;
  double *st = state + (int64_t) k * ncol + ic;
  const int64_t fs = (int64_t) Nmax * ncol;

  // units, rh/rhf1d/pyrh_compute1dray.c:230-270
  const double CUBE_CM = CM_TO_M * CM_TO_M * CM_TO_M;
  double T = atm[Nmax + k];
  double ne = atm[2 * Nmax + k] / CUBE_CM;
  double vz = atm[3 * Nmax + k] * KM_TO_M;
  double vturb = atm[4 * Nmax + k] * KM_TO_M;
  double B = atm[5 * Nmax + k] / 1e4;
  double nHtot = atm[8 * Nmax + k] / CUBE_CM;
  double gam = atm[6 * Nmax + k], chB = atm[7 * Nmax + k];
  double cosg, cos2x, sin2x;
  if (mu == 1.0) {
    cosg = cos(gam);
    cos2x = cos(2.0 * chB);
    sin2x = sin(2.0 * chB);
  } else {  // rh/rhf1d/project.c:59-76
    double mux = sqrt(1.0 - mu * mu), muy = 0.0, csc_theta = 1.0 / sqrt(1.0 - mu * mu);
    double sin_gamma = sin(gam), bx = sin_gamma * cos(chB), by = sin_gamma * sin(chB),
           bz = cos(gam), b3 = mux * bx + muy * by + mu * bz, b1 = csc_theta * (bz - mu * b3),
           b2 = csc_theta * (muy * bx - mux * by);
    cosg = b3;
    cos2x = (b1 * b1 - b2 * b2) / (1.0 - b3 * b3);
    sin2x = 2.0 * b1 * b2 / (1.0 - b3 * b3);
  }
  st[RHG_CF_T * fs] = T;
  st[RHG_CF_NE * fs] = ne;
  st[RHG_CF_VLOS * fs] = vz;
  st[RHG_CF_VTURB * fs] = vturb;
  st[RHG_CF_B * fs] = B;
  st[RHG_CF_COSG * fs] = cosg;
  st[RHG_CF_COS2X * fs] = cos2x;
  st[RHG_CF_SIN2X * fs] = sin2x;
  st[RHG_CF_NHTOT * fs] = nHtot;

  // Saha-Boltzmann normalisation per atom, rh/ltepops.c:33-113
  const double c1 = (HPLANCK / (2.0 * PI * M_ELECTRON)) * (HPLANCK / KBOLTZMANN);
  const double cNe_T = 0.5 * ne * pow(c1 / T, 1.5);
  // the atoms that are molecule nuclei: total LTE normalisation and the
  // neutral fraction (rh/chemequil.c:222-236)
  double nstar0n[NN], fn0[NN];
#pragma unroll
  for (int i = 0; i < NN; i++) {
    const int ia = ldc(M.nucl_atom)[i], l0 = ldc(M.atom_level_ptr)[ia], nl = ldc(M.atom_level_ptr)[ia + 1] - l0;
    nstar0n[i] = (ldc(M.atom_abund)[ia] * nHtot) / atom_sum(M, l0, nl, T, cNe_T);
    double ntot = ldc(M.atom_abund)[ia] * nHtot, f = 0.0;
    for (int j = 0; j < nl; j++) {
      if (ldc(M.level_stage)[l0 + j] > 0) break;
      f += (j == 0) ? nstar0n[i] : level_ratio(M, l0, j, T, cNe_T) * nstar0n[i];
    }
    fn0[i] = f / ntot;
  }
  double nnuc[NN], PhiHmin;
  extern __shared__ double lds_mol[];
  const int lstride = blockDim.x;
  double *lds_phi = lds_mol + threadIdx.x;
  double *lds_nmol = lds_phi + M.nmol * lstride;
  chem_eq<NN>(M, T, ne, nHtot, fn0, nnuc, lds_phi, lds_nmol, nullptr, lstride, PhiHmin);
  double fracn[NN];
#pragma unroll
  for (int i = 0; i < NN; i++) fracn[i] = nnuc[i] / (ldc(M.atom_abund)[ldc(M.nucl_atom)[i]] * nHtot);
  const double nHmin = ne * (nnuc[0] * PhiHmin);
  st[RHG_CF_NHMIN * fs] = nHmin;
  st[RHG_CF_NH2 * fs] = M.mol_h2 >= 0 ? lds_nmol[M.mol_h2 * lstride] : 0.0;
  st[RHG_CF_NOH * fs] = M.mol_oh >= 0 ? lds_nmol[M.mol_oh * lstride] : 0.0;
  st[RHG_CF_NCH * fs] = M.mol_ch >= 0 ? lds_nmol[M.mol_ch * lstride] : 0.0;
  // hydrogen is atom 0 and nucleus 0
  const int lH = ldc(M.atom_level_ptr)[0], nlH = ldc(M.atom_level_ptr)[1] - lH;
  const double nH0 = nstar0n[0] * fracn[0];
  const double np = (level_ratio(M, lH, nlH - 1, T, cNe_T) * nstar0n[0]) * fracn[0];
  st[RHG_CF_NH0 * fs] = nH0;
  st[RHG_CF_NP * fs] = np;

  // bound-free continua: lower-level population n_i.  In LTE n = n*, so the
  // emissivity factor n_j n_i*/n_j* of rh/metal.c:120-140 (n_p n_i*/n_p* for
  // hydrogen, rh/hydrogen.c:300-330, whose continuum level is the proton
  // level) is n_i as well.  The continua come grouped by atom; each atom's
  // normalisation is formed once.
  {
    int cur = -1;
    double nst0 = 0.0, fr = 1.0;
    for (int ci = 0; ci < M.ncont; ci++) {
      const int ia = ldc(M.cont_atom)[ci], l0 = ldc(M.atom_level_ptr)[ia];
      if (ia != cur) {
        cur = ia;
        int inuc = -1;
#pragma unroll
        for (int i = 0; i < NN; i++)
          if (ldc(M.nucl_atom)[i] == ia) inuc = i;
        if (inuc >= 0) {
#pragma unroll
          for (int i = 0; i < NN; i++)
            if (i == inuc) { nst0 = nstar0n[i]; fr = fracn[i]; }
        } else {
          nst0 = (ldc(M.atom_abund)[ia] * nHtot) / atom_sum(M, l0, ldc(M.atom_level_ptr)[ia + 1] - l0, T, cNe_T);
          fr = 1.0;
        }
      }
      const int i = ldc(M.cont_i)[ci];
      st[cont_field(ci) * fs] = ((i == 0) ? nst0 : level_ratio(M, l0, i, T, cNe_T) * nst0) * fr;
    }
  }

  // Kurucz lines: lambda-independent per-depth quantities,
  // rh/kurucz.c:662-677 and rh/kurucz.c:743-775
  const int lbase = line_field_base(M.ncont);
  // what depends on T (and n_e) alone, formed once for all lines: the
  // partition-function bracket and the Saha factor
  const Linear1At pfT(M.npf, M.Tpf, T);
  const double C1 = (HPLANCK / (2.0 * PI * M_ELECTRON)) * (HPLANCK / KBOLTZMANN);
  const double CT_ne = M.nline > 0 ? 2.0 * pow(C1 / T, -1.5) / ne : 0.0;
  for (int l = 0; l < M.nline; l++) {
    int ie = ldc(M.line_el)[l], stg = ldc(M.line_stage)[l];
    // stage populations of the element, rh/ltepops.c:116-164
    const int npf = M.npf, ns = ldc(M.el_nstage)[ie];
    const double *lp = M.el_logpf + ldc(M.el_pf_ptr)[ie];
    CDouble *ip = ldc(M.el_ionpot) + ldc(M.el_pf_ptr)[ie];
    double sum = 1.0, Uk, Ukp1, nst = 1.0, nstage = 1.0;
    Uk = pfT(lp);
    for (int i = 1; i < ns; i++) {
      Ukp1 = pfT(lp + i * npf);
      nst = nst * CT_ne * exp(Ukp1 - Uk - ip[i - 1] / (KBOLTZMANN * T));
      sum += nst;
      if (i == stg) nstage = nst;
      Uk = Ukp1;
    }
    double n0 = ldc(M.el_abund)[ie] * nHtot / sum;
    nstage = (stg == 0) ? n0 : nstage * n0;
    double pf = pfT(lp + stg * npf);
    double kT = 1.0 / (KBOLTZMANN * T);
    double ni_gi = nstage * exp(-ldc(M.line_Ei)[l] * kT - pf);
    double nj_gj = ni_gi * exp(-ldc(M.line_hc_la)[l] * kT);
    double chi_l = ldc(M.line_Bijhc4pi)[l] * (ni_gi - nj_gj);
    double eta_l = ldc(M.line_Bijhc4pi)[l] * ldc(M.line_twohnu3_c2)[l] * nj_gj;
    double vtherm = 2.0 * KBOLTZMANN / (AMU * ldc(M.el_weight)[ie]);
    double vbroad = sqrt(vtherm * T + vturb * vturb);
    double adamp = 0.0;
    if (ldc(M.line_Grad)[l] != 0.0) {
      double GvdW;
      switch (ldc(M.line_vdw)[l]) {
        case RHG_VDW_UNSOLD: GvdW = ldc(M.line_cross)[l] * pow(T, 0.3); break;
        case RHG_VDW_BARKLEM: GvdW = ldc(M.line_cross)[l] * pow(T, (1.0 - ldc(M.line_alpha)[l]) / 2.0); break;
        default: GvdW = ldc(M.line_GvdW)[l]; break;
      }
      adamp = (ldc(M.line_Grad)[l] + ldc(M.line_GStark)[l] * ne + GvdW * (nHtot - np)) *
              (ldc(M.line_lambda0)[l] * NM_TO_M) / (4.0 * PI * vbroad);
    }
    st[(lbase + 4 * l) * fs] = chi_l;
    st[(lbase + 4 * l + 1) * fs] = eta_l;
    st[(lbase + 4 * l + 2) * fs] = adamp;
    st[(lbase + 4 * l + 3) * fs] = 1.0 / vbroad;
  }

  // background continuum at the reference wavelength (for convertScales;
  // k_geometry adds line opacity there if the reference wavelength has any)
  {
    double chi_ai, eta_ai;
    continuum_point(M, M.ref_index, st, fs, chi_ai, eta_ai);
    st[RHG_CF_HEIGHT * fs] = chi_ai;
  }
  // depth scale input
  {
    double sc = atm[k];
    double *tau = st + (int64_t) (nfield - 1) * fs;
    if (atm_scale == 0) *tau = pow10d(sc);
    else if (atm_scale == 1) *tau = pow10d(sc) * (G_TO_KG / (CM_TO_M * CM_TO_M));
    else *tau = sc * KM_TO_M;
  }
}

// ---------------------------------------------------------------------------
// Stage 2: convertScales, rh/rhf1d/multiatmos.c:100-177
__global__ void k_geometry(const DevModel *__restrict__ Mp, const int32_t *__restrict__ ndep, int ncol, int Nmax,
                           int atm_scale, double mu, double *__restrict__ state, int nfield,
                           int32_t *__restrict__ moving_out, double *__restrict__ chiref_out) {
  const DevModel &M = *Mp;
  const int ic = blockIdx.x * blockDim.x + threadIdx.x;
  if (ic >= ncol) return;
  const int N = ndep[ic];
  const int64_t fs = (int64_t) Nmax * ncol;
  double *st0 = state + ic;
#define H(k) st0[(int64_t) RHG_CF_HEIGHT * fs + (int64_t) (k) * ncol]   // chi_ref on entry
#define SCALE(k) st0[(int64_t) (nfield - 1) * fs + (int64_t) (k) * ncol]
#define NH(k) st0[(int64_t) RHG_CF_NHTOT * fs + (int64_t) (k) * ncol]
  const double rho_f = AMU * M.wght_per_H;
  // moving atmosphere flag, rh/rhf1d/pyrh_compute1dray.c:272-278
  int moving = 0;
  for (int k = 0; k < N; k++)
    if (fabs(st0[(int64_t) RHG_CF_VLOS * fs + (int64_t) k * ncol]) >= M.vmacro_tresh) {
      moving = 1;
      break;
    }
  moving_out[ic] = moving;
  if (ldc(M.lam_line_ptr)[M.ref_index + 1] > ldc(M.lam_line_ptr)[M.ref_index]) {
    for (int k = 0; k < N; k++) {
      double lchi, leta, q[3], e[3];
      line_point<double>(M, M.ref_index, st0 + (int64_t) k * ncol, fs, mu, lchi, q, leta, e, moving);
      H(k) += lchi;
    }
  }
  if (chiref_out)
    for (int k = 0; k < N; k++) chiref_out[(int64_t) ic * Nmax + k] = H(k);
  if (atm_scale == 0) {
    double chi_prev = H(0), tau_prev = SCALE(0), hp = 0.0;
    H(0) = 0.0;
    for (int k = 1; k < N; k++) {
      double chik = H(k);
      double hk = hp - 2.0 * (SCALE(k) - tau_prev) / (chi_prev + chik);
      H(k) = hk;
      hp = hk;
      chi_prev = chik;
      tau_prev = SCALE(k);
    }
    // h_zero = Linear(tau, height, 1.0) (bisection, rh/linear.c)
    double hz;
    if (1.0 <= SCALE(0)) hz = H(0);
    else if (1.0 >= SCALE(N - 1)) hz = H(N - 1);
    else {
      int lo = 0, hi = N;
      while (hi - lo > 1) {
        int idx = (hi + lo) >> 1;
        if (1.0 >= SCALE(idx)) lo = idx; else hi = idx;
      }
      const int j = lo;
      double fx = (SCALE(j + 1) - 1.0) / (SCALE(j + 1) - SCALE(j));
      hz = fx * H(j) + (1 - fx) * H(j + 1);
    }
    for (int k = 0; k < N; k++) H(k) = H(k) - hz;
  } else if (atm_scale == 1) {
    double chi_prev = H(0), hp = 0.0;
    double taup = chi_prev / (rho_f * NH(0)) * SCALE(0);
    double tauv[RHG_MAX_DEPTH];
    tauv[0] = taup;
    H(0) = 0.0;
    for (int k = 1; k < N; k++) {
      double chik = H(k);
      double hk = hp - 2.0 * (SCALE(k) - SCALE(k - 1)) / (rho_f * NH(k - 1) + rho_f * NH(k));
      taup = taup + 0.5 * (chi_prev + chik) * (hp - hk);
      tauv[k] = taup;
      H(k) = hk;
      hp = hk;
      chi_prev = chik;
    }
    double hz;
    if (1.0 <= tauv[0]) hz = H(0);
    else if (1.0 >= tauv[N - 1]) hz = H(N - 1);
    else {
      int j = locate_asc(N, tauv, 1.0);
      double fx = (tauv[j + 1] - 1.0) / (tauv[j + 1] - tauv[j]);
      hz = fx * H(j) + (1 - fx) * H(j + 1);
    }
    for (int k = 0; k < N; k++) H(k) = H(k) - hz;
  } else {
    for (int k = 0; k < N; k++) H(k) = SCALE(k);
  }
  // geometric path lengths of the depth intervals
  double *dz = st0 + (int64_t) RHG_CF_DZ * fs;
  dz[0] = 0.0;
  for (int k = 1; k < N; k++) dz[(int64_t) k * ncol] = fabs(H(k - 1) - H(k));
#undef H
#undef SCALE
#undef NH
}

// ---------------------------------------------------------------------------
// pyrh's get_scales (rh/rhf1d/pyrh_hse.c:402-553): convertScales
// (rh/rhf1d/multiatmos.c:100-177) with all three depth scales kept, from the
// column state k_state left (the background opacity at the reference
// wavelength in the height field).  One thread per column; tau, height [m]
// and cmass [kg m^-2] are [ncol][Nmax], zero past ndep.  The input scale is
// returned converted to SI (the reference leaves that output unset).
__global__ void k_scales(const DevModel *__restrict__ Mp, const int32_t *__restrict__ ndep,
                         int ncol, int Nmax, int atm_scale, const double *__restrict__ state,
                         int nfield, double *__restrict__ tau, double *__restrict__ height,
                         double *__restrict__ cmass) {
  const DevModel &M = *Mp;
  const int ic = blockIdx.x * blockDim.x + threadIdx.x;
  if (ic >= ncol) return;
  const int N = ndep[ic];
  const int64_t fs = (int64_t) Nmax * ncol;
  const double *st0 = state + ic;
  auto CHI = [&](int k) { return st0[(int64_t) RHG_CF_HEIGHT * fs + (int64_t) k * ncol]; };
  auto SC = [&](int k) { return st0[(int64_t) (nfield - 1) * fs + (int64_t) k * ncol]; };
  auto NHT = [&](int k) { return st0[(int64_t) RHG_CF_NHTOT * fs + (int64_t) k * ncol]; };
  const double rho_f = AMU * M.wght_per_H;
  double *tc = tau + (int64_t) ic * Nmax, *hc = height + (int64_t) ic * Nmax,
         *mc = cmass + (int64_t) ic * Nmax;
  for (int k = 0; k < N; k++) {
    if (atm_scale == 0) tc[k] = SC(k);
    else if (atm_scale == 1) mc[k] = SC(k);
    else hc[k] = SC(k);
  }
  if (atm_scale == 1) {
    hc[0] = 0.0;
    tc[0] = CHI(0) / (rho_f * NHT(0)) * mc[0];
    for (int k = 1; k < N; k++) {
      hc[k] = hc[k - 1] - 2.0 * (mc[k] - mc[k - 1]) / (rho_f * NHT(k - 1) + rho_f * NHT(k));
      tc[k] = tc[k - 1] + 0.5 * (CHI(k - 1) + CHI(k)) * (hc[k - 1] - hc[k]);
    }
  } else if (atm_scale == 0) {
    hc[0] = 0.0;
    mc[0] = (tc[0] / CHI(0)) * (rho_f * NHT(0));
    for (int k = 1; k < N; k++) {
      hc[k] = hc[k - 1] - 2.0 * (tc[k] - tc[k - 1]) / (CHI(k - 1) + CHI(k));
      mc[k] = mc[k - 1] + 0.5 * (rho_f * NHT(k - 1) + rho_f * NHT(k)) * (hc[k - 1] - hc[k]);
    }
  } else if (N > 1) {
    const double T0 = st0[(int64_t) RHG_CF_T * fs], ne0 = st0[(int64_t) RHG_CF_NE * fs];
    mc[0] = (NHT(0) * M.totalAbund + ne0) * (KBOLTZMANN * T0 / M.gravity);
    tc[0] = 0.5 * CHI(0) * (hc[0] - hc[1]);
    if (tc[0] > 1.0) tc[0] = 0.0;
    for (int k = 1; k < N; k++) {
      mc[k] = mc[k - 1] + 0.5 * (rho_f * NHT(k - 1) + rho_f * NHT(k)) * (hc[k - 1] - hc[k]);
      tc[k] = tc[k - 1] + 0.5 * (CHI(k - 1) + CHI(k)) * (hc[k - 1] - hc[k]);
    }
  }
  if (atm_scale == 0 || atm_scale == 1) {
    // h_zero = Linear(tau, height, 1.0), rh/linear.c:22-54
    double hz;
    if (1.0 <= tc[0]) hz = hc[0];
    else if (1.0 >= tc[N - 1]) hz = hc[N - 1];
    else {
      const int j = locate_asc(N, tc, 1.0);
      const double fx = (tc[j + 1] - 1.0) / (tc[j + 1] - tc[j]);
      hz = fx * hc[j] + (1 - fx) * hc[j + 1];
    }
    for (int k = 0; k < N; k++) hc[k] = hc[k] - hz;
  }
  for (int k = N; k < Nmax; k++) tc[k] = hc[k] = mc[k] = 0.0;
}

// ---------------------------------------------------------------------------
// Stage 3: continuum opacity and emissivity for every output wavelength at
// one (column, depth), the terms of continuum_point() with what depends on
// T and ne alone evaluated once.  The bound-free sums
// sum_c alpha_c(lambda) n_i,c are accumulated BF_TILE wavelengths at a time
// in registers (one load of n_i,c per tile; the tile's cross-sections are
// wave-uniform) and handed to the per-wavelength pass through LDS.  One
// exponential per wavelength: exp(-hc/(k lambda T)) is the reciprocal of the
// Planck function's exp(hc/(k lambda T)).
// Output layout: cchi[(io * Nmax + k) * ncol + column].  With wavelength
// grids (blk_grid != nullptr, ncol a multiple of 64) the columns of block
// ic / 64 belong to grid g = blk_grid[ic / 64] (-1: none) and take the
// bound-free tiles gtile[2 g] .. gtile[2 g + 1] - 1 only, the tiles that hold
// the grid's wavelengths, and in them the wavelengths io whose bit is set in
// the grid's mask, bit io % 32 of gmask[g * gwords + io / 32] (the GRID
// instantiation; the other ignores the three grid arguments).
template <typename R, bool GRID>
__global__ void __launch_bounds__(CONT_BLOCK)
k_continuum(const DevModel *__restrict__ Mp, const int32_t *__restrict__ ndep, int ncol, int Nmax,
            const double *__restrict__ state, int nfield, const int32_t *__restrict__ out_ls,
            int nout, R *__restrict__ cchi, R *__restrict__ ceta,
            const int32_t *__restrict__ blk_grid, const int32_t *__restrict__ gtile,
            const int32_t *__restrict__ gmask, int gwords) {
  const DevModel &M = *Mp;
  __shared__ double bf_lds[BF_TILE][CONT_BLOCK];
  const int64_t gid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (int64_t) ncol * Nmax) return;
  const int k = (int) (gid / ncol), ic = (int) (gid % ncol);
  if (k >= ndep[ic]) return;
  int tile0 = 0, tile1 = M.ntile;
  const int32_t *gm = nullptr;
  if constexpr (GRID) {
    // a wavefront is 64 consecutive slots of one depth (ncol is a multiple
    // of 64), one block: its grid and tile range are wave-uniform, and kept
    // in scalar registers so that the tile tables stay scalar loads
    const int g = __builtin_amdgcn_readfirstlane(blk_grid[ic >> 6]);
    if (g < 0) return;
    tile0 = __builtin_amdgcn_readfirstlane(gtile[2 * g]);
    tile1 = __builtin_amdgcn_readfirstlane(gtile[2 * g + 1]);
    gm = gmask + (int64_t) g * gwords;
  }
  const int64_t fs = (int64_t) Nmax * ncol;
  const double *st = state + (int64_t) k * ncol + ic;
  const double T = st[RHG_CF_T * fs], ne = st[RHG_CF_NE * fs];
  const double nHmin = st[RHG_CF_NHMIN * fs], nH0 = st[RHG_CF_NH0 * fs], np = st[RHG_CF_NP * fs];
  const double nH2 = st[RHG_CF_NH2 * fs], nOH = st[RHG_CF_NOH * fs], nCH = st[RHG_CF_NCH * fs];
  const double C0 = Q_ELECTRON * Q_ELECTRON / (4.0 * PI * EPSILON_0) / sqrt(M_ELECTRON);
  const double sigma_ff = 4.0 / 3.0 * sqrt(2.0 * PI / (3.0 * KBOLTZMANN)) * (C0 * C0 * C0) /
                          (HPLANCK * CLIGHT);
  // T- and ne-only factors (rh/hydrogen.c:509-523 table indices)
  const double iT = 1.0 / T;
  const double pe = ne * KBOLTZMANN * T;
  const double ti_hmin = clamp_index(NTHETA, M.hminff_theta, THETA0 * iT);
  const double ti_h2m = clamp_index(NTHETA_H2, M.h2mff_theta, THETA0 * iT);
  const double ti_h2p = clamp_index(NTEMP_H2P, M.h2pff_temp, T);
  const double sqrt_theta = sqrt(THETA0 * iT), sff_T = sigma_ff / sqrt(T);
  const double hff_fac = ne * np;
  const double y_fac = T * (1.0 / (HPLANCK * CLIGHT));
  const bool oh_in = !(T < ldc(M.oh_T)[0] || T > ldc(M.oh_T)[NTOH - 1]);
  const bool ch_in = !(T < ldc(M.ch_T)[0] || T > ldc(M.ch_T)[NTCH - 1]);
  double ti_oh = 0.0, ti_ch = 0.0;
  if (oh_in) {
    int i2 = locate_asc(NTOH, M.oh_T, T);
    ti_oh = (i2 == NTOH - 1) ? (double) i2 : (double) i2 + (T - ldc(M.oh_T)[i2]) / (ldc(M.oh_T)[i2 + 1] - ldc(M.oh_T)[i2]);
  }
  if (ch_in) {
    int i2 = locate_asc(NTCH, M.ch_T, T);
    ti_ch = (i2 == NTCH - 1) ? (double) i2 : (double) i2 + (T - ldc(M.ch_T)[i2]) / (ldc(M.ch_T)[i2 + 1] - ldc(M.ch_T)[i2]);
  }

  for (int t = tile0; t < tile1; t++) {
    const int io0 = t * BF_TILE;
    {
      double bf[BF_TILE];
#pragma unroll
      for (int j = 0; j < BF_TILE; j++) bf[j] = 0.0;
      const int e0 = ldc(M.tile_ptr)[t], e1 = ldc(M.tile_ptr)[t + 1];
      for (int e = e0; e < e1; e++) {
        const double ni = st[cont_field(ldc(M.tile_cont)[e]) * fs];
        const CDouble *al = ldc(M.tile_alpha) + (int64_t) e * BF_TILE;
#pragma unroll
        for (int j = 0; j < BF_TILE; j++) bf[j] += al[j] * ni;
      }
#pragma unroll
      for (int j = 0; j < BF_TILE; j++) bf_lds[j][threadIdx.x] = bf[j];
    }
    const int nj = (nout - io0 < BF_TILE) ? nout - io0 : BF_TILE;
#pragma unroll 1
    for (int j = 0; j < nj; j++) {
      const int io = io0 + j;
      if constexpr (GRID)
        if (!(((uint32_t) ldc(gm)[io >> 5] >> (io & 31)) & 1u)) continue;
      const double bfj = bf_lds[j][threadIdx.x];
      const CDouble *lc = ldc(M.lamconst) + (int64_t) ldc(out_ls)[io] * RHG_LC_NCONST;
      const double hc_kla = lc[RHG_LC_HC_KLA], twohnu3 = lc[RHG_LC_TWOHNU3_C2];
      const double x = hc_kla * iT;
      // Planck function, rh/planck.c:38-63, and the stimulated-emission
      // factor exp(-x); past MAX_EXPONENT the Planck function is 0 and
      // exp(-x) < 1e-65 is taken as 0
      const bool in_range = x <= MAX_EXPONENT;
      const double ex = exp(in_range ? x : MAX_EXPONENT);
      const double Bnu = in_range ? twohnu3 / (ex - 1.0) : 0.0;
      const double stim = in_range ? 1.0 / ex : 0.0;
      double chi_ai = 0.0, eta_ai = 0.0, chi;
      if (lc[RHG_LC_HMIN_BF_ON] != 0.0) {
        chi_ai += nHmin * (1.0 - stim) * lc[RHG_LC_HMIN_BF_ALPHA];
        eta_ai += nHmin * twohnu3 * stim * lc[RHG_LC_HMIN_BF_ALPHA];
      }
      if (lc[RHG_LC_HMIN_FF_LIDX] >= 0.0) {
        double kappa = bilinear(NTHETA, NFF, M.hminff_kappa, ti_hmin, lc[RHG_LC_HMIN_FF_LIDX]);
        chi = (nH0 * 1.0E-29) * pe * kappa;
      } else {
        double Ck = (KBOLTZMANN * THETA0 * 1.0E-32), theta_n = 1.0;
        chi = 0.0;
        for (int n = 1; n < 6; n++) {
          theta_n *= sqrt_theta;
          chi += theta_n * lc[RHG_LC_HFF_LONG0 + n - 1];
        }
        chi *= nH0 * (ne * Ck);
      }
      chi_ai += chi;
      eta_ai += chi * Bnu;
      chi_ai *= lc[RHG_LC_HMIN_FUDGE];
      eta_ai *= lc[RHG_LC_HMIN_FUDGE];
      if (lc[RHG_LC_OH_EIDX] >= 0.0 && oh_in) {
        double kappa = exp(LG10 * bilinear(NTOH, NEOH, M.oh_cross, ti_oh, lc[RHG_LC_OH_EIDX])) *
                       (CM_TO_M * CM_TO_M);
        chi_ai += nOH * (1.0 - stim) * kappa;
        eta_ai += nOH * twohnu3 * stim * kappa;
      }
      if (lc[RHG_LC_CH_EIDX] >= 0.0 && ch_in) {
        double kappa = exp(LG10 * bilinear(NTCH, NECH, M.ch_cross, ti_ch, lc[RHG_LC_CH_EIDX])) *
                       (CM_TO_M * CM_TO_M);
        chi_ai += nCH * (1.0 - stim) * kappa;
        eta_ai += nCH * twohnu3 * stim * kappa;
      }
      // bound-free (hydrogen and metals), rh/hydrogen.c:300-330, rh/metal.c:120-140
      {
        const double expla = (lc[RHG_LC_BF_HC_KLA] == hc_kla) ? stim : exp(-lc[RHG_LC_BF_HC_KLA] * iT);
        chi_ai += (1.0 - expla) * bfj;
        eta_ai += lc[RHG_LC_BF_TWOHNU3_C2] * expla * bfj;
      }
      {
        double y = lc[RHG_LC_GFF_CY] * y_fac, x3 = lc[RHG_LC_GFF_X3];
        double gIII = 1.0 + 0.1728 * x3 * (1.0 + y) - 0.0496 * (x3 * x3) * (1.0 + (1.0 + y) * 0.33333333 * y);
        double g_ff = (gIII > 1.0) ? gIII : 1.0;
        chi = sff_T * lc[RHG_LC_HFF_NU3] * hff_fac * (1.0 - stim) * g_ff;
        chi_ai += chi;
        eta_ai += chi * Bnu;
      }
      if (lc[RHG_LC_H2P_LIDX] >= 0.0) {
        double kappa = bilinear(NTEMP_H2P, NFF_H2P, M.h2pff_kappa, ti_h2p, lc[RHG_LC_H2P_LIDX]);
        chi = (nH0 * 1.0E-29) * (np * 1.0E-20) * kappa;
        chi_ai += chi;
        eta_ai += chi * Bnu;
      }
      if (lc[RHG_LC_H2M_LIDX] >= 0.0 && nH2 > 0.0) {
        double kappa = bilinear(NTHETA_H2, NFF_H2, M.h2mff_kappa, ti_h2m, lc[RHG_LC_H2M_LIDX]);
        chi = (nH2 * 1.0E-29) * pe * kappa;
        chi_ai += chi;
        eta_ai += chi * Bnu;
      }
      const int64_t o = ((int64_t) io * Nmax + k) * ncol + ic;
      cchi[o] = (R) chi_ai;
      ceta[o] = (R) eta_ai;
    }
  }
}

#define RHG_KSTATE(NN)                                                                       \
  template __global__ void k_state<NN>(const DevModel *, const double *, const int32_t *,        \
                                       const int32_t *, int, int, double, int, double *, int);
RHG_KSTATE(1)
RHG_KSTATE(2)
RHG_KSTATE(3)
RHG_KSTATE(4)
RHG_KSTATE(5)
RHG_KSTATE(6)
RHG_KSTATE(7)
RHG_KSTATE(8)
#undef RHG_KSTATE

#define RHG_KCONT(R, G)                                                                       \
  template __global__ void k_continuum<R, G>(const DevModel *, const int32_t *, int, int,      \
                                             const double *, int, const int32_t *, int, R *, R *, \
                                             const int32_t *, const int32_t *, const int32_t *, int);
RHG_KCONT(double, false)
RHG_KCONT(double, true)
RHG_KCONT(float, false)
RHG_KCONT(float, true)
#undef RHG_KCONT

// ---------------------------------------------------------------------------
// Column reordering.  perm[i] is the caller's column at sorted position i.
__global__ void k_gather_cols(const double *__restrict__ src, double *__restrict__ dst,
                              const int32_t *__restrict__ perm, int ncol, int64_t len) {
  const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t) ncol * len) return;
  const int64_t i = p / len, j = p % len;
  dst[p] = src[(int64_t) perm[i] * len + j];
}

__global__ void k_scatter_cols(const double *__restrict__ src, double *__restrict__ dst,
                               const int32_t *__restrict__ perm, int ncol, int64_t len) {
  const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t) ncol * len) return;
  const int64_t i = p / len, j = p % len;
  dst[(int64_t) perm[i] * len + j] = src[p];
}

template <typename T>
__global__ void k_scatter_last(const T *__restrict__ src, T *__restrict__ dst,
                               const int32_t *__restrict__ perm, int ncol, int64_t nouter) {
  const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t) ncol * nouter) return;
  const int64_t o = p / ncol;
  const int i = (int) (p % ncol);
  dst[o * ncol + perm[i]] = src[p];
}
template __global__ void k_scatter_last<float>(const float *, float *, const int32_t *, int, int64_t);
template __global__ void k_scatter_last<double>(const double *, double *, const int32_t *, int,
                                                int64_t);

}  // namespace rhg
