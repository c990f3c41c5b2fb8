// gfx950 kernels for batched LTE Stokes synthesis (FP64 path).
//
// Stage 1  k_state      one thread per (column, depth): unit conversion,
//                       B projection, Saha-Boltzmann populations of every
//                       model atom, chemical equilibrium (Newton), Kurucz
//                       element populations, per-line opacity constants
//                       and the background opacity at the reference
//                       wavelength.  Writes the column state.
// Stage 2  k_geometry   one thread per column: tau500 -> height
//                       (convertScales, a serial scan of Ndep).
// Stage 3  k_continuum  one thread per (column, depth, wavelength):
//                       continuum opacity and emissivity.
// Stage 4  k_formal     one thread per (column, wavelength): line opacity
//                       for the emergent ray computed depth by depth while
//                       the DELO-Bezier / Bezier / Feautrier recursion runs
//                       in registers.
#include "rhg_kernels.hpp"

namespace rhg {

// ---------------------------------------------------------------------------
// Stage 1: column state

// Saha-Boltzmann ratio of level i to level 0 (before normalisation),
// rh/ltepops.c:87-100
__device__ __forceinline__ double level_ratio(const DevModel &M, int l0, int i, double T,
                                              double cNe_T) {
  double dE = M.level_E[l0 + i] - M.level_E[l0];
  double gi0 = M.level_g[l0 + i] / M.level_g[l0];
  int dZ = M.level_stage[l0 + i] - M.level_stage[l0];
  double dE_kT = dE / (KBOLTZMANN * T);
  double r = gi0 * exp(-dE_kT);
  for (int m = 1; m <= dZ; m++) r /= cNe_T;
  return r;
}

__device__ double equilconstant(const DevModel &M, int im, double T) {
  int neqc = M.mol_neqc[im];
  const double *c = M.mol_eqc + 8 * im;
  double kT, t, eqc = 0.0, theta, cgs_to_SI = 1.0;
  if (T < M.mol_tmin[im] || T > M.mol_tmax[im]) return 0.0;
  switch (M.mol_fit[im]) {
    case 0:
    case 1: {
      t = (M.mol_fit[im] == 1) ? T * 1.0E-4 : T;
      kT = KBOLTZMANN * T;
      int mk = M.mol_nnuclei[im] - 1 - M.mol_charge[im];
      eqc = c[0];
      for (int i = 1; i < neqc; i++) eqc = eqc * t + c[i];
      eqc = exp(M.mol_ediss[im] / kT + eqc - 1.5 * mk * log(T));
      cgs_to_SI = pow(CM_TO_M * CM_TO_M * CM_TO_M, (double) mk);
      break;
    }
    case 2:
    case 3:
      theta = THETA0 / T;
      t = log10(theta);
      kT = KBOLTZMANN * T;
      eqc = c[0];
      for (int i = 1; i < neqc; i++) eqc = eqc * t + c[i];
      eqc = pow10d((M.mol_ediss[im] / EV) * theta - eqc) * kT;
      cgs_to_SI = 1.0;
      break;
    case 4:
      theta = THETA0 / T;
      kT = KBOLTZMANN * T;
      eqc = c[0];
      for (int i = 1; i < neqc; i++) eqc = eqc * theta + c[i];
      eqc = kT * kT * pow10d(-eqc);
      cgs_to_SI = pow(CM_TO_M * CM_TO_M * CM_TO_M / ERG_TO_JOULE, (double) (M.mol_nnuclei[im] - 1));
      break;
  }
  return eqc * cgs_to_SI;
}

// Solve the 4x4 (Nn x Nn) system by Gaussian elimination with partial
// pivoting; a[Nn][NNUC_MAX], b[Nn] overwritten with the solution.
__device__ void small_solve(int n, double a[NNUC_MAX][NNUC_MAX], double *b) {
  for (int j = 0; j < n; j++) {
    int p = j;
    double big = fabs(a[j][j]);
    for (int i = j + 1; i < n; i++)
      if (fabs(a[i][j]) > big) { big = fabs(a[i][j]); p = i; }
    if (p != j) {
      for (int k = 0; k < n; k++) { double t = a[j][k]; a[j][k] = a[p][k]; a[p][k] = t; }
      double t = b[j]; b[j] = b[p]; b[p] = t;
    }
    double inv = 1.0 / a[j][j];
    for (int i = j + 1; i < n; i++) {
      double f = a[i][j] * inv;
      for (int k = j; k < n; k++) a[i][k] -= f * a[j][k];
      b[i] -= f * b[j];
    }
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = b[i];
    for (int k = i + 1; k < n; k++) s -= a[i][k] * b[k];
    b[i] = s / a[i][i];
  }
}

// Chemical equilibrium at one depth (rh/chemequil.c:107-393).  The Newton
// system [[D, B], [C, I]] (nuclei, molecules) is reduced to the nuclei
// block by the Schur complement S = D - B C, solved, and corrected once
// with the residual of the full system, as the reference refines its LU
// solution once (rh/ludcmp.c:55-84).
__device__ void chem_eq(const DevModel &M, double T, double ne, double nHtot, const double *fn0,
                        double *nnuc, double *nmol, double &PhiHmin_out) {
  const int Nn = M.nnuclei, Nm = M.nmol;
  const double CI = (HPLANCK / (2.0 * PI * M_ELECTRON)) * (HPLANCK / KBOLTZMANN);
  double a[NNUC_MAX], Phi[NMOL_MAX], prev_n[NNUC_MAX + NMOL_MAX];
  for (int i = 0; i < Nn; i++) a[i] = M.nucl_abund[i] * nHtot;
  double PhiHmin = 0.25 * pow(CI / T, 1.5) * exp(0.754 * EV / (KBOLTZMANN * T));
  double fHmin = ne * fn0[0] * PhiHmin;
  for (int i = 0; i < Nm; i++) Phi[i] = equilconstant(M, i, T);
  for (int i = 0; i < Nn; i++) nnuc[i] = a[i];
  for (int i = 0; i < Nm; i++) nmol[i] = 0.0;
  for (int i = 0; i < Nn; i++) prev_n[i] = nnuc[i];
  for (int i = 0; i < Nm; i++) prev_n[Nn + i] = 0.0;

  for (int niter = 1; niter <= N_MAX_CHEM_ITER; niter++) {
    double f1[NNUC_MAX], f2[NMOL_MAX], saha[NMOL_MAX];
    for (int i = 0; i < Nn; i++) f1[i] = nnuc[i] - a[i];
    for (int i = 0; i < Nm; i++) f2[i] = nmol[i] - 0.0;
    f1[0] += fHmin * nnuc[0];
    for (int i = 0; i < Nm; i++) {
      double s = Phi[i];
      int nel = M.mol_nel[i];
      for (int j = 0; j < nel; j++) {
        int nu = M.mol_el_nucl[4 * i + j];
        s *= pow(fn0[nu] * nnuc[nu], (double) M.mol_el_count[4 * i + j]);
        f1[nu] += M.mol_el_count[4 * i + j] * nmol[i];
      }
      s /= pow(ne, (double) M.mol_charge[i]);
      f2[i] -= s;
      saha[i] = s;
    }
    // S = D - B C, rhs = f1 - B f2  (C[m][nu] = -saha_m count / n_nu)
    double S[NNUC_MAX][NNUC_MAX], S0[NNUC_MAX][NNUC_MAX], r[NNUC_MAX], x1[NNUC_MAX];
    for (int i = 0; i < Nn; i++) {
      for (int j = 0; j < Nn; j++) S[i][j] = (i == j) ? 1.0 : 0.0;
      r[i] = f1[i];
    }
    S[0][0] += fHmin;
    for (int m = 0; m < Nm; m++) {
      int nel = M.mol_nel[m];
      for (int j = 0; j < nel; j++) {
        int nu = M.mol_el_nucl[4 * m + j];
        double cnt = M.mol_el_count[4 * m + j];
        r[nu] -= cnt * f2[m];
        for (int jj = 0; jj < nel; jj++) {
          int nu2 = M.mol_el_nucl[4 * m + jj];
          double C = -saha[m] * (M.mol_el_count[4 * m + jj] / nnuc[nu2]);
          S[nu][nu2] -= cnt * C;
        }
      }
    }
    for (int i = 0; i < Nn; i++) {
      x1[i] = r[i];
      for (int j = 0; j < Nn; j++) S0[i][j] = S[i][j];
    }
    small_solve(Nn, S, x1);
    // one refinement step on the reduced system
    {
      double res[NNUC_MAX];
      for (int i = 0; i < Nn; i++) {
        res[i] = r[i];
        for (int j = 0; j < Nn; j++) res[i] -= S0[i][j] * x1[j];
      }
      double S1[NNUC_MAX][NNUC_MAX];
      for (int i = 0; i < Nn; i++)
        for (int j = 0; j < Nn; j++) S1[i][j] = S0[i][j];
      small_solve(Nn, S1, res);
      for (int i = 0; i < Nn; i++) x1[i] += res[i];
    }
    double dnmax = 0.0;
    for (int m = 0; m < Nm; m++) {
      double x2 = f2[m];
      int nel = M.mol_nel[m];
      for (int j = 0; j < nel; j++) {
        int nu = M.mol_el_nucl[4 * m + j];
        double C = -saha[m] * (M.mol_el_count[4 * m + j] / nnuc[nu]);
        x2 -= C * x1[nu];
      }
      nmol[m] -= x2;
    }
    for (int i = 0; i < Nn; i++) nnuc[i] -= x1[i];
    // MaxChange over the last two solutions (rh/maxchange.c:32-50)
    for (int i = 0; i < Nn; i++) {
      if (nnuc[i]) dnmax = fmax(dnmax, fabs((nnuc[i] - prev_n[i]) / nnuc[i]));
      prev_n[i] = nnuc[i];
    }
    for (int i = 0; i < Nm; i++) {
      if (nmol[i]) dnmax = fmax(dnmax, fabs((nmol[i] - prev_n[Nn + i]) / nmol[i]));
      prev_n[Nn + i] = nmol[i];
    }
    if (dnmax <= CHEM_ITER_LIMIT) break;
  }
  PhiHmin_out = PhiHmin;
}

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
  if (lc[RHG_LC_OH_EIDX] >= 0.0) {
    if (T < M.oh_T[0] || T > M.oh_T[NTOH - 1]) { chi = 0.0; eta = 0.0; }
    else {
      int i2 = locate_asc(NTOH, M.oh_T, T);
      double ti = (i2 == NTOH - 1) ? (double) i2
                                   : (double) i2 + (T - M.oh_T[i2]) / (M.oh_T[i2 + 1] - M.oh_T[i2]);
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
    if (T < M.ch_T[0] || T > M.ch_T[NTCH - 1]) { chi = 0.0; eta = 0.0; }
    else {
      int i2 = locate_asc(NTCH, M.ch_T, T);
      double ti = (i2 == NTCH - 1) ? (double) i2
                                   : (double) i2 + (T - M.ch_T[i2]) / (M.ch_T[i2 + 1] - M.ch_T[i2]);
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
  const int cp0 = M.lam_cont_ptr[ls], cp1 = M.lam_cont_ptr[ls + 1];
  double expla = exp(-lc[RHG_LC_BF_HC_KLA] / T);
  {
    bool any = false;
    chi = eta = 0.0;
    for (int e = cp0; e < cp1; e++) {
      int ci = M.lam_cont_idx[e];
      if (M.cont_atom[ci] != 0) continue;
      any = true;
      double sigma = M.lam_cont_alpha[e];
      double ni = st[(RHG_CF_NFIXED + 3 * ci) * fs], r = st[(RHG_CF_NFIXED + 3 * ci + 1) * fs];
      double gijk = r * expla;
      chi += sigma * (1.0 - expla) * ni;
      eta += lc[RHG_LC_BF_TWOHNU3_C2] * gijk * sigma * np;
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
    int ci = M.lam_cont_idx[e];
    if (M.cont_atom[ci] == 0) continue;
    double alpha_la = M.lam_cont_alpha[e];
    double ni = st[(RHG_CF_NFIXED + 3 * ci) * fs], r = st[(RHG_CF_NFIXED + 3 * ci + 1) * fs],
           nj = st[(RHG_CF_NFIXED + 3 * ci + 2) * fs];
    double gijk = r * expla;
    chi += alpha_la * (1.0 - expla) * ni;
    eta += lc[RHG_LC_BF_TWOHNU3_C2] * gijk * alpha_la * nj;
  }
  chi_ai += chi * 1.0;
  eta_ai += eta * 1.0;
}

__global__ void k_state(DevModel M, const double *__restrict__ atmos, const int32_t *__restrict__ ndep,
                        int ncol, int Nmax, double mu, int atm_scale, double *__restrict__ state,
                        int nfield) {
  int64_t gid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (int64_t) ncol * Nmax) return;
  const int ic = (int) (gid / Nmax), k = (int) (gid % Nmax);
  if (k >= ndep[ic]) return;
  const double *atm = atmos + (int64_t) ic * 9 * Nmax;
  double *st = state + (int64_t) ic * nfield * Nmax + k;
  const int64_t fs = Nmax;

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
  double nstar0[RHG_MAX_ATOM], fn0[NNUC_MAX], frac[RHG_MAX_ATOM];
  for (int ia = 0; ia < M.natom; ia++) {
    int l0 = M.atom_level_ptr[ia], nl = M.atom_level_ptr[ia + 1] - l0;
    double sum = 1.0;
    for (int i = 1; i < nl; i++) sum += level_ratio(M, l0, i, T, cNe_T);
    nstar0[ia] = (M.atom_abund[ia] * nHtot) / sum;
    frac[ia] = 1.0;
  }
  // neutral fractions of the nuclei (rh/chemequil.c:222-236)
  for (int i = 0; i < M.nnuclei; i++) {
    int ia = M.nucl_atom[i], l0 = M.atom_level_ptr[ia], nl = M.atom_level_ptr[ia + 1] - l0;
    double ntot = M.atom_abund[ia] * nHtot, f = 0.0;
    for (int j = 0; j < nl; j++) {
      if (M.level_stage[l0 + j] > 0) break;
      f += (j == 0) ? nstar0[ia] : level_ratio(M, l0, j, T, cNe_T) * nstar0[ia];
    }
    fn0[i] = f / ntot;
  }
  double nnuc[NNUC_MAX], nmol[NMOL_MAX], PhiHmin;
  chem_eq(M, T, ne, nHtot, fn0, nnuc, nmol, PhiHmin);
  for (int i = 0; i < M.nnuclei; i++) {
    int ia = M.nucl_atom[i];
    frac[ia] = nnuc[i] / (M.atom_abund[ia] * nHtot);
  }
  const double nHmin = ne * (nnuc[0] * PhiHmin);
  st[RHG_CF_NHMIN * fs] = nHmin;
  st[RHG_CF_NH2 * fs] = M.mol_h2 >= 0 ? nmol[M.mol_h2] : 0.0;
  st[RHG_CF_NOH * fs] = M.mol_oh >= 0 ? nmol[M.mol_oh] : 0.0;
  st[RHG_CF_NCH * fs] = M.mol_ch >= 0 ? nmol[M.mol_ch] : 0.0;
  const int lH = M.atom_level_ptr[0], nlH = M.atom_level_ptr[1] - lH;
  const double nH0 = nstar0[0] * frac[0];
  const double np = (level_ratio(M, lH, nlH - 1, T, cNe_T) * nstar0[0]) * frac[0];
  st[RHG_CF_NH0 * fs] = nH0;
  st[RHG_CF_NP * fs] = np;

  // bound-free continuum level terms: n_i, n_i*/n_j*, n_j
  for (int ci = 0; ci < M.ncont; ci++) {
    int ia = M.cont_atom[ci], l0 = M.atom_level_ptr[ia];
    int i = M.cont_i[ci], j = M.cont_j[ci];
    double ni = ((i == 0) ? nstar0[ia] : level_ratio(M, l0, i, T, cNe_T) * nstar0[ia]) * frac[ia];
    double nj = ((j == 0) ? nstar0[ia] : level_ratio(M, l0, j, T, cNe_T) * nstar0[ia]) * frac[ia];
    st[(RHG_CF_NFIXED + 3 * ci) * fs] = ni;
    st[(RHG_CF_NFIXED + 3 * ci + 1) * fs] = ni / nj;
    st[(RHG_CF_NFIXED + 3 * ci + 2) * fs] = nj;
  }

  // Kurucz lines: lambda-independent per-depth quantities,
  // rh/kurucz.c:662-677 and rh/kurucz.c:743-775
  const int lbase = RHG_CF_NFIXED + 3 * M.ncont;
  for (int l = 0; l < M.nline; l++) {
    int ie = M.line_el[l], stg = M.line_stage[l];
    // stage populations of the element, rh/ltepops.c:116-164
    const double C1 = (HPLANCK / (2.0 * PI * M_ELECTRON)) * (HPLANCK / KBOLTZMANN);
    const int npf = M.npf, ns = M.el_nstage[ie];
    const double *lp = M.el_logpf + M.el_pf_ptr[ie];
    const double *ip = M.el_ionpot + M.el_pf_ptr[ie];
    double CT_ne = 2.0 * pow(C1 / T, -1.5) / ne, sum = 1.0, Uk, Ukp1, nst = 1.0, nstage = 1.0;
    Uk = linear1(npf, M.Tpf, lp, T);
    for (int i = 1; i < ns; i++) {
      Ukp1 = linear1(npf, M.Tpf, lp + i * npf, T);
      nst = nst * CT_ne * exp(Ukp1 - Uk - ip[i - 1] / (KBOLTZMANN * T));
      sum += nst;
      if (i == stg) nstage = nst;
      Uk = Ukp1;
    }
    double n0 = M.el_abund[ie] * nHtot / sum;
    nstage = (stg == 0) ? n0 : nstage * n0;
    double pf = linear1(npf, M.Tpf, lp + stg * npf, T);
    double kT = 1.0 / (KBOLTZMANN * T);
    double ni_gi = nstage * exp(-M.line_Ei[l] * kT - pf);
    double nj_gj = ni_gi * exp(-M.line_hc_la[l] * kT);
    double chi_l = M.line_Bijhc4pi[l] * (ni_gi - nj_gj);
    double eta_l = M.line_Bijhc4pi[l] * M.line_twohnu3_c2[l] * nj_gj;
    double vtherm = 2.0 * KBOLTZMANN / (AMU * M.el_weight[ie]);
    double vbroad = sqrt(vtherm * T + vturb * vturb);
    double adamp = 0.0;
    if (M.line_Grad[l] != 0.0) {
      double GvdW;
      switch (M.line_vdw[l]) {
        case RHG_VDW_UNSOLD: GvdW = M.line_cross[l] * pow(T, 0.3); break;
        case RHG_VDW_BARKLEM: GvdW = M.line_cross[l] * pow(T, (1.0 - M.line_alpha[l]) / 2.0); break;
        default: GvdW = M.line_GvdW[l]; break;
      }
      adamp = (M.line_Grad[l] + M.line_GStark[l] * ne + GvdW * (nHtot - np)) *
              (M.line_lambda0[l] * NM_TO_M) / (4.0 * PI * vbroad);
    }
    st[(lbase + 4 * l) * fs] = chi_l;
    st[(lbase + 4 * l + 1) * fs] = eta_l;
    st[(lbase + 4 * l + 2) * fs] = adamp;
    st[(lbase + 4 * l + 3) * fs] = vbroad;
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
__global__ void k_geometry(DevModel M, const int32_t *__restrict__ ndep, int ncol, int Nmax,
                           int atm_scale, double mu, double *__restrict__ state, int nfield,
                           int32_t *__restrict__ moving_out, double *__restrict__ chiref_out) {
  int ic = blockIdx.x * blockDim.x + threadIdx.x;
  if (ic >= ncol) return;
  const int N = ndep[ic];
  double *st = state + (int64_t) ic * nfield * Nmax;
  double *h = st + (int64_t) RHG_CF_HEIGHT * Nmax;      // holds chi_ref on entry
  const double *scale = st + (int64_t) (nfield - 1) * Nmax;
  const double *nH = st + (int64_t) RHG_CF_NHTOT * Nmax;
  const double rho_f = AMU * M.wght_per_H;
  // moving atmosphere flag, rh/rhf1d/pyrh_compute1dray.c:272-278
  int moving = 0;
  for (int k = 0; k < N; k++)
    if (fabs(st[(int64_t) RHG_CF_VLOS * Nmax + k]) >= M.vmacro_tresh) { moving = 1; break; }
  moving_out[ic] = moving;
  if (M.lam_line_ptr[M.ref_index + 1] > M.lam_line_ptr[M.ref_index]) {
    for (int k = 0; k < N; k++) {
      double lchi, leta, q[3], e[3];
      line_point(M, M.ref_index, st + k, Nmax, mu, lchi, q, leta, e, nullptr, moving);
      h[k] += lchi;
    }
  }
  if (chiref_out)
    for (int k = 0; k < N; k++) chiref_out[(int64_t) ic * Nmax + k] = h[k];
  if (atm_scale == 0) {
    double chi_prev = h[0], tau_prev = scale[0], hp = 0.0;
    h[0] = 0.0;
    for (int k = 1; k < N; k++) {
      double chik = h[k];
      double hk = hp - 2.0 * (scale[k] - tau_prev) / (chi_prev + chik);
      h[k] = hk;
      hp = hk;
      chi_prev = chik;
      tau_prev = scale[k];
    }
    // h_zero = Linear(tau, height, 1.0) (bisection, rh/linear.c)
    double hz;
    if (1.0 <= scale[0]) hz = h[0];
    else if (1.0 >= scale[N - 1]) hz = h[N - 1];
    else {
      int j = locate_asc(N, scale, 1.0);
      double fx = (scale[j + 1] - 1.0) / (scale[j + 1] - scale[j]);
      hz = fx * h[j] + (1 - fx) * h[j + 1];
    }
    for (int k = 0; k < N; k++) h[k] = h[k] - hz;
  } else if (atm_scale == 1) {
    double chi_prev = h[0], hp = 0.0;
    double taup = chi_prev / (rho_f * nH[0]) * scale[0];
    double tauv[RHG_MAX_DEPTH];
    tauv[0] = taup;
    h[0] = 0.0;
    for (int k = 1; k < N; k++) {
      double chik = h[k];
      double hk = hp - 2.0 * (scale[k] - scale[k - 1]) / (rho_f * nH[k - 1] + rho_f * nH[k]);
      taup = taup + 0.5 * (chi_prev + chik) * (hp - hk);
      tauv[k] = taup;
      h[k] = hk;
      hp = hk;
      chi_prev = chik;
    }
    double hz;
    if (1.0 <= tauv[0]) hz = h[0];
    else if (1.0 >= tauv[N - 1]) hz = h[N - 1];
    else {
      int j = locate_asc(N, tauv, 1.0);
      double fx = (tauv[j + 1] - 1.0) / (tauv[j + 1] - tauv[j]);
      hz = fx * h[j] + (1 - fx) * h[j + 1];
    }
    for (int k = 0; k < N; k++) h[k] = h[k] - hz;
  } else {
    for (int k = 0; k < N; k++) h[k] = scale[k];
  }
}

// ---------------------------------------------------------------------------
// Stage 3: continuum per (column, depth, output wavelength)
__global__ void k_continuum(DevModel M, const int32_t *__restrict__ ndep, int ncol, int Nmax,
                            const double *__restrict__ state, int nfield,
                            const int32_t *__restrict__ out_ls, int nout,
                            double *__restrict__ cchi, double *__restrict__ ceta) {
  // block: x = wavelength tile, y = depth; grid.z = column
  const int io = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  const int ic = blockIdx.z;
  if (io >= nout || k >= ndep[ic]) return;
  const double *st = state + (int64_t) ic * nfield * Nmax + k;
  double chi, eta;
  continuum_point(M, out_ls[io], st, Nmax, chi, eta);
  int64_t o = ((int64_t) ic * Nmax + k) * nout + io;
  cchi[o] = chi;
  ceta[o] = eta;
}

}  // namespace rhg
