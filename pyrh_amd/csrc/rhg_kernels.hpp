// Kernel declarations, the device copy of the model, and the per-depth line
// opacity shared by the stages.
#pragma once

#include "rhg_device.hpp"

#define RHG_MAX_ATOM 32
#define RHG_MAX_DEPTH 512

namespace rhg {

// Device mirror of rhg_model_t: the same fields with device pointers
// (filled by the host from the caller's model, see rhg_api.hip).
struct DevModel {
  int stokes_mode, magneto_optical;
  double vmacro_tresh, vmicro_char, wght_per_H, totalAbund, gravity;
  int nspect, ref_index;
  const double *lamconst;
  const int32_t *solver, *lam_line_ptr, *lam_line_idx, *lam_cont_ptr, *lam_cont_idx;
  const double *lam_cont_alpha;
  int natom;
  const int32_t *atom_level_ptr;
  const double *level_E, *level_g;
  const int32_t *level_stage;
  const double *atom_abund;
  int ncont;
  const int32_t *cont_atom, *cont_i, *cont_j;
  int nnuclei;
  const int32_t *nucl_atom, *nucl_elem;
  const double *nucl_abund;
  int nmol;
  const int32_t *mol_fit, *mol_charge, *mol_nnuclei;
  const double *mol_ediss, *mol_tmin, *mol_tmax;
  const int32_t *mol_neqc;
  const double *mol_eqc;
  const int32_t *mol_nel, *mol_el_nucl, *mol_el_count;
  int mol_h2, mol_oh, mol_ch;
  int npf;
  const double *Tpf;
  int nel;
  const int32_t *el_nstage, *el_pf_ptr;
  const double *el_logpf, *el_ionpot, *el_abund, *el_weight;
  int nline;
  const int32_t *line_el, *line_stage;
  const double *line_lambda0, *line_Ei, *line_hc_la, *line_Bijhc4pi, *line_twohnu3_c2,
      *line_Grad, *line_GStark, *line_GvdW;
  const int32_t *line_vdw;
  const double *line_cross, *line_alpha, *line_larmor;
  const int32_t *line_polar, *line_comp_ptr, *comp_q;
  const double *comp_shift, *comp_strength;
  const double *hminff_theta, *hminff_kappa, *h2mff_theta, *h2mff_kappa, *h2pff_temp,
      *h2pff_kappa, *oh_T, *oh_cross, *ch_T, *ch_cross;
};

// index of the first per-line field in the column state
__device__ __forceinline__ int line_field_base(const DevModel &M) {
  return RHG_CF_NFIXED + 3 * M.ncont;
}

// Kurucz-line opacity and emissivity of the emergent ray (to_obs = 1) at
// wavelength ls and the depth that st points at (field stride fs),
// rh/kurucz.c:600-715 and RLKProfile rh/kurucz.c:729-827.
// chiI/etaI accumulate from zero; chiQ[3]/etaQ[3] likewise.
__device__ __forceinline__ void line_point(const DevModel &M, int ls, const double *st, int64_t fs,
                                           double mu, double &chiI, double *chiQ, double &etaI,
                                           double *etaQ, double * /*unused*/, int moving) {
  const double lambda = M.lamconst[(int64_t) ls * RHG_LC_NCONST + RHG_LC_LAMBDA];
  const int p0 = M.lam_line_ptr[ls], p1 = M.lam_line_ptr[ls + 1];
  const int lb = line_field_base(M);
  chiI = 0.0; etaI = 0.0;
  chiQ[0] = chiQ[1] = chiQ[2] = 0.0;
  etaQ[0] = etaQ[1] = etaQ[2] = 0.0;
  const double vz = st[RHG_CF_VLOS * fs];
  for (int e = p0; e < p1; e++) {
    const int l = M.lam_line_idx[e];
    const double chi_l = st[(lb + 4 * l) * fs], eta_l = st[(lb + 4 * l + 1) * fs];
    const double adamp = st[(lb + 4 * l + 2) * fs], vbroad = st[(lb + 4 * l + 3) * fs];
    double v = (lambda / M.line_lambda0[l] - 1.0) * CLIGHT / vbroad;
    if (moving) v += (mu * vz) / vbroad;
    const double sv = 1.0 / (SQRTPI * vbroad);
    double phi, phi_Q = 0.0, phi_U = 0.0, phi_V = 0.0;
    const bool polar = M.line_polar[l] != 0;
    if (M.line_Grad[l] != 0.0) {
      if (polar) {
        const double cosg = st[RHG_CF_COSG * fs];
        const double sin2_gamma = 1.0 - cosg * cosg;
        const double vB = M.line_larmor[l] * st[RHG_CF_B * fs] / vbroad;
        const double sign = 1.0;
        double phi_sm = 0.0, phi_pi = 0.0, phi_sp = 0.0, H, F;
        const int c0 = M.line_comp_ptr[l], c1 = M.line_comp_ptr[l + 1];
        for (int nz = c0; nz < c1; nz++) {
          H = voigt_humlicek(adamp, v - M.comp_shift[nz] * vB, &F);
          const int q = M.comp_q[nz];
          if (q == -1) phi_sm += M.comp_strength[nz] * H;
          else if (q == 0) phi_pi += M.comp_strength[nz] * H;
          else phi_sp += M.comp_strength[nz] * H;
        }
        const double phi_sigma = phi_sp + phi_sm;
        const double phi_delta = 0.5 * phi_pi - 0.25 * phi_sigma;
        phi = (phi_delta * sin2_gamma + 0.5 * phi_sigma) * sv;
        phi_Q = sign * phi_delta * sin2_gamma * st[RHG_CF_COS2X * fs] * sv;
        phi_U = phi_delta * sin2_gamma * st[RHG_CF_SIN2X * fs] * sv;
        phi_V = sign * 0.5 * (phi_sp - phi_sm) * cosg * sv;
      } else {
        phi = voigt_armstrong(adamp, v) * sv;
      }
    } else {
      phi = (fabs(v) <= MAX_GAUSS_DOPPLER) ? exp(-v * v) : 0.0;
      phi = phi * sv;
    }
    if (phi != 0.0) {
      chiI += chi_l * phi;
      etaI += eta_l * phi;
      if (polar && M.line_Grad[l] != 0.0) {
        chiQ[0] += chi_l * phi_Q;
        chiQ[1] += chi_l * phi_U;
        chiQ[2] += chi_l * phi_V;
        etaQ[0] += eta_l * phi_Q;
        etaQ[1] += eta_l * phi_U;
        etaQ[2] += eta_l * phi_V;
      }
    }
  }
}

// Work census of one emergent ray (column, wavelength) over depth indices
// 1..N-1: the same line window, Doppler offsets and Voigt branches as
// line_point(), with counts in place of arithmetic (slots: RHG_WORK_*).
__device__ __forceinline__ void line_census(const DevModel &M, int ls, const double *st, int64_t fs,
                                            double mu, int moving, int32_t *cnt) {
  const double lambda = M.lamconst[(int64_t) ls * RHG_LC_NCONST + RHG_LC_LAMBDA];
  const int p0 = M.lam_line_ptr[ls], p1 = M.lam_line_ptr[ls + 1];
  const int lb = line_field_base(M);
  const double vz = st[RHG_CF_VLOS * fs];
  for (int e = p0; e < p1; e++) {
    const int l = M.lam_line_idx[e];
    const double adamp = st[(lb + 4 * l + 2) * fs], vbroad = st[(lb + 4 * l + 3) * fs];
    double v = (lambda / M.line_lambda0[l] - 1.0) * CLIGHT / vbroad;
    if (moving) v += (mu * vz) / vbroad;
    cnt[RHG_WORK_LINE]++;
    if (M.line_Grad[l] != 0.0) {
      if (M.line_polar[l] != 0) {
        const double vB = M.line_larmor[l] * st[RHG_CF_B * fs] / vbroad;
        for (int nz = M.line_comp_ptr[l]; nz < M.line_comp_ptr[l + 1]; nz++)
          cnt[RHG_WORK_VOIGT_I + humlicek_region(adamp, v - M.comp_shift[nz] * vB)]++;
      } else {
        cnt[RHG_WORK_ARMSTRONG]++;
      }
    } else {
      cnt[RHG_WORK_GAUSS]++;
    }
  }
}

__global__ void k_state(DevModel M, const double *atmos, const int32_t *ndep, int ncol, int Nmax,
                        double mu, int atm_scale, double *state, int nfield);
__global__ void k_geometry(DevModel M, const int32_t *ndep, int ncol, int Nmax, int atm_scale,
                           double mu, double *state, int nfield, int32_t *moving_out,
                           double *chiref_out);
__global__ void k_continuum(DevModel M, const int32_t *ndep, int ncol, int Nmax,
                            const double *state, int nfield, const int32_t *out_ls, int nout,
                            double *cchi, double *ceta);
__global__ void k_formal(DevModel M, const int32_t *ndep, const int32_t *moving, int ncol,
                         int Nmax, const double *state, int nfield, const int32_t *out_ls,
                         int nout, const double *cchi, const double *ceta, double mu,
                         double *stokes);

__global__ void k_census(DevModel M, const int32_t *ndep, const int32_t *moving, int ncol,
                         int Nmax, const double *state, int nfield, const int32_t *out_ls,
                         int nout, double mu, int32_t *census);
__global__ void k_census_sum(const int32_t *census, int ncol, int nout, int64_t *col_sum);

}  // namespace rhg
