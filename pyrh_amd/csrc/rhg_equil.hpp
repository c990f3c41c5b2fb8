// Chemical equilibrium at one depth, shared by the column-state kernel of
// the synthesis and the hydrostatic-equilibrium builder: molecular
// equilibrium constants (rh/chemequil.c:452-533) and the Newton solution of
// the nuclei and molecule densities (rh/chemequil.c:107-393).
#pragma once

#include "rhg_kernels.hpp"

namespace rhg {

__device__ inline double equilconstant(const DevModel &M, int im, double T) {
  int neqc = ldc(M.mol_neqc)[im];
  CDouble *c = ldc(M.mol_eqc) + 8 * im;
  double kT, t, eqc = 0.0, theta, cgs_to_SI = 1.0;
  if (T < ldc(M.mol_tmin)[im] || T > ldc(M.mol_tmax)[im]) return 0.0;
  switch (ldc(M.mol_fit)[im]) {
    case 0:
    case 1: {
      t = (ldc(M.mol_fit)[im] == 1) ? T * 1.0E-4 : T;
      kT = KBOLTZMANN * T;
      int mk = ldc(M.mol_nnuclei)[im] - 1 - ldc(M.mol_charge)[im];
      eqc = c[0];
      for (int i = 1; i < neqc; i++) eqc = eqc * t + c[i];
      eqc = exp(ldc(M.mol_ediss)[im] / kT + eqc - 1.5 * mk * log(T));
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
      eqc = pow10d((ldc(M.mol_ediss)[im] / EV) * theta - eqc) * kT;
      cgs_to_SI = 1.0;
      break;
    case 4:
      theta = THETA0 / T;
      kT = KBOLTZMANN * T;
      eqc = c[0];
      for (int i = 1; i < neqc; i++) eqc = eqc * theta + c[i];
      eqc = kT * kT * pow10d(-eqc);
      cgs_to_SI = pow(CM_TO_M * CM_TO_M * CM_TO_M / ERG_TO_JOULE,
                      (double) (ldc(M.mol_nnuclei)[im] - 1));
      break;
  }
  return eqc * cgs_to_SI;
}

// x^n for a small integer n (the stoichiometric counts and charges of the
// molecules), by multiplication
__device__ __forceinline__ double ipow(double x, int n) {
  if (n == 0) return 1.0;
  double r = x;
  for (int i = 1; i < (n < 0 ? -n : n); i++) r *= x;
  return n < 0 ? 1.0 / r : r;
}

// Solve the NN x NN system a x = b by Gaussian elimination with partial
// pivoting; b is overwritten with the solution.  NN is a compile-time size,
// so the arrays live in registers.
template <int NN>
__device__ __forceinline__ void small_solve(double (&a)[NN][NN], double (&b)[NN]) {
#pragma unroll
  for (int j = 0; j < NN; j++) {
    int p = j;
    double big = fabs(a[j][j]);
#pragma unroll
    for (int i = j + 1; i < NN; i++)
      if (fabs(a[i][j]) > big) { big = fabs(a[i][j]); p = i; }
    // swap rows j and p (p >= j), with compile-time indices
#pragma unroll
    for (int i = j + 1; i < NN; i++) {
      if (i == p) {
#pragma unroll
        for (int k = 0; k < NN; k++) { double t = a[j][k]; a[j][k] = a[i][k]; a[i][k] = t; }
        double t = b[j]; b[j] = b[i]; b[i] = t;
      }
    }
    double inv = 1.0 / a[j][j];
#pragma unroll
    for (int i = j + 1; i < NN; i++) {
      double f = a[i][j] * inv;
#pragma unroll
      for (int k = j; k < NN; k++) a[i][k] -= f * a[j][k];
      b[i] -= f * b[j];
    }
  }
#pragma unroll
  for (int i = NN - 1; i >= 0; i--) {
    double s = b[i];
#pragma unroll
    for (int k = i + 1; k < NN; k++) s -= a[i][k] * b[k];
    b[i] = s / a[i][i];
  }
}

// Chemical equilibrium at one depth (rh/chemequil.c:107-393) for NN nuclei.
// The Newton system [[D, B], [C, I]] (nuclei, molecules) is reduced to the
// nuclei block by the Schur complement S = D - B C, solved, and corrected
// once with the residual of the reduced system, as the reference refines
// its LU solution once (rh/ludcmp.c:55-84).  Molecule stoichiometry is the
// dense table M.mol_cnt[m][nucleus]; the per-molecule equilibrium constant,
// density and Saha product live in LDS (lds_phi, lds_nmol, lds_saha, stride
// blockDim.x), everything per nucleus in registers.  With lds_saha nullptr
// the Saha products are formed again where they are reused, by the same
// operations (so the same values), which saves a third of the LDS.
template <int NN>
__device__ void chem_eq(const DevModel &M, double T, double ne, double nHtot,
                        const double (&fn0)[NN], double (&nnuc)[NN], double *lds_phi,
                        double *lds_nmol, double *lds_saha, int ls_stride, double &PhiHmin_out,
                        bool phi_ready = false, const double *tcache = nullptr) {
  const int Nm = M.nmol;
  const double CI = (HPLANCK / (2.0 * PI * M_ELECTRON)) * (HPLANCK / KBOLTZMANN);
  double a[NN];
#pragma unroll
  for (int i = 0; i < NN; i++) a[i] = ldc(M.nucl_abund)[i] * nHtot;
  // tcache: pow(CI/T, 1.5) and exp(E_ION_HMIN/kT) formed by a caller that
  // solves this temperature again
  double PhiHmin = tcache ? 0.25 * tcache[0] * tcache[1]
                          : 0.25 * pow(CI / T, 1.5) * exp(0.754 * EV / (KBOLTZMANN * T));
  double fHmin = ne * fn0[0] * PhiHmin;
  // phi_ready: lds_phi already holds the constants at this T (a caller
  // that solves the same temperature again)
  for (int m = 0; m < Nm; m++) {
    if (!phi_ready) lds_phi[m * ls_stride] = equilconstant(M, m, T);
    lds_nmol[m * ls_stride] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < NN; i++) nnuc[i] = a[i];

  for (int niter = 1; niter <= N_MAX_CHEM_ITER; niter++) {
    double f1[NN], S[NN][NN], r[NN], x1[NN];
    // the Jacobian's count / n_nu (rh/chemequil.c:300-304) as count * (1 / n_nu):
    // one division per nucleus instead of one per (molecule, nucleus), a
    // last-place difference
    double inv_n[NN];
#pragma unroll
    for (int i = 0; i < NN; i++) inv_n[i] = 1.0 / nnuc[i];
#pragma unroll
    for (int i = 0; i < NN; i++) f1[i] = nnuc[i] - a[i];
    f1[0] += fHmin * nnuc[0];
    double fn[NN];
#pragma unroll
    for (int i = 0; i < NN; i++) fn[i] = fn0[i] * nnuc[i];
    // Saha product of molecule m at this iteration's nucleus densities
    auto saha = [&](int m) {
      CDouble *cnt = ldc(M.mol_cnt) + m * NNUC_MAX;
      double sv = lds_phi[m * ls_stride];
#pragma unroll
      for (int i = 0; i < NN; i++)
        if (cnt[i] != 0.0) sv *= ipow(fn[i], (int) cnt[i]);
      sv /= ipow(ne, ldc(M.mol_charge)[m]);
      return sv;
    };
    auto saha_of = [&](int m) { return lds_saha ? lds_saha[m * ls_stride] : saha(m); };
    for (int m = 0; m < Nm; m++) {
      CDouble *cnt = ldc(M.mol_cnt) + m * NNUC_MAX;
      const double nm = lds_nmol[m * ls_stride];
#pragma unroll
      for (int i = 0; i < NN; i++)
        if (cnt[i] != 0.0) f1[i] += cnt[i] * nm;
      if (lds_saha) lds_saha[m * ls_stride] = saha(m);
    }
    // S = D - B C, rhs = f1 - B f2  (C[m][nu] = -saha_m count_nu / n_nu,
    // f2[m] = n_m - saha_m)
#pragma unroll
    for (int i = 0; i < NN; i++) {
#pragma unroll
      for (int j = 0; j < NN; j++) S[i][j] = (i == j) ? 1.0 : 0.0;
      r[i] = f1[i];
    }
    S[0][0] += fHmin;
    for (int m = 0; m < Nm; m++) {
      CDouble *cnt = ldc(M.mol_cnt) + m * NNUC_MAX;
      const double sv = saha_of(m);
      const double f2 = lds_nmol[m * ls_stride] - sv;
      double C[NN];
#pragma unroll
      for (int j = 0; j < NN; j++) C[j] = -sv * (cnt[j] * inv_n[j]);
#pragma unroll
      for (int i = 0; i < NN; i++)
        if (cnt[i] != 0.0) {
          r[i] -= cnt[i] * f2;
#pragma unroll
          for (int j = 0; j < NN; j++)
            if (cnt[j] != 0.0) S[i][j] -= cnt[i] * C[j];
        }
    }
    double S0[NN][NN], S1[NN][NN];
#pragma unroll
    for (int i = 0; i < NN; i++) {
      x1[i] = r[i];
#pragma unroll
      for (int j = 0; j < NN; j++) { S0[i][j] = S[i][j]; S1[i][j] = S[i][j]; }
    }
    small_solve<NN>(S, x1);
    {
      double res[NN];
#pragma unroll
      for (int i = 0; i < NN; i++) {
        res[i] = r[i];
#pragma unroll
        for (int j = 0; j < NN; j++) res[i] -= S0[i][j] * x1[j];
      }
      small_solve<NN>(S1, res);
#pragma unroll
      for (int i = 0; i < NN; i++) x1[i] += res[i];
    }
    // back substitution for the molecules and MaxChange over the last two
    // solutions (rh/maxchange.c:32-50)
    double dnmax = 0.0;
    for (int m = 0; m < Nm; m++) {
      CDouble *cnt = ldc(M.mol_cnt) + m * NNUC_MAX;
      const double sv = saha_of(m);
      const double nold = lds_nmol[m * ls_stride];
      double x2 = nold - sv;
#pragma unroll
      for (int j = 0; j < NN; j++)
        if (cnt[j] != 0.0) x2 -= (-sv * (cnt[j] * inv_n[j])) * x1[j];
      const double nnew = nold - x2;
      lds_nmol[m * ls_stride] = nnew;
      if (nnew) dnmax = fmax(dnmax, fabs((nnew - nold) / nnew));
    }
#pragma unroll
    for (int i = 0; i < NN; i++) {
      const double nold = nnuc[i];
      nnuc[i] -= x1[i];
      if (nnuc[i]) dnmax = fmax(dnmax, fabs((nnuc[i] - nold) / nnuc[i]));
    }
    if (dnmax <= CHEM_ITER_LIMIT) break;
  }
  PhiHmin_out = PhiHmin;
}

}  // namespace rhg
