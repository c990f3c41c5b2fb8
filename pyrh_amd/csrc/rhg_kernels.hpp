// Kernel declarations, the device copy of the model, and the per-depth line
// opacity shared by the stages.
#pragma once

#include "rhg_device.hpp"
#include "rhg_dual.hpp"

#define RHG_MAX_ATOM 32
#define RHG_MAX_DEPTH 512

namespace rhg {

// Device mirror of rhg_model_t: the same fields with device pointers
// (filled by the host from the caller's model, see rhg_api.hip).
// Read-only tables through the constant address space: a load with a
// wave-uniform index becomes a scalar load (through the scalar cache)
// instead of a vector memory access per lane.  Only for data no kernel
// writes while it runs.
typedef __attribute__((address_space(4))) const double CDouble;
typedef __attribute__((address_space(4))) const int32_t CInt;
__device__ __forceinline__ CDouble *ldc(const double *p) { return (CDouble *) p; }
__device__ __forceinline__ CInt *ldc(const int32_t *p) { return (CInt *) p; }

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
  // derived on the device copy (rhg_open):
  // (lambda/lambda0 - 1) c for every (wavelength, line) entry of lam_line_idx
  const double *lam_line_dv;
  // bound-free cross-sections by tiles of BF_TILE output wavelengths: tile t
  // holds the continua active anywhere in it, entries tile_ptr[t]..
  // tile_ptr[t+1]-1, with cross-sections tile_alpha[entry][BF_TILE] (zero
  // where a continuum is inactive)
  int ntile;
  const int32_t *tile_ptr, *tile_cont;
  const double *tile_alpha;
  // molecule stoichiometry as a dense table [nmol][NNUC_MAX] of counts
  const double *mol_cnt;
};

// output wavelengths per bound-free tile of k_continuum, and its block size
constexpr int BF_TILE = 16;
constexpr int CONT_BLOCK = 256;

// Workgroup -> (column block of 64, output wavelength) for the per-ray
// kernels.  Workgroups are dispatched round-robin over the 8 XCDs (linear
// id mod 8), each XCD with its own L2; the mapping sends all nout
// wavelengths of a column block to one XCD, consecutively, so that the
// block's column state is fetched into that L2 once and reused by every
// wavelength.  Launch xcd_grid(ncb, nout) workgroups; ids that map past the
// last column block return cb >= ncb.
constexpr int NXCD = 8;
__host__ __device__ __forceinline__ int64_t xcd_grid(int ncb, int nout) {
  return (int64_t) ((ncb + NXCD - 1) / NXCD) * NXCD * nout;
}
__device__ __forceinline__ void xcd_tile(int64_t b, int nout, int &cb, int &io) {
  const int64_t per_group = (int64_t) NXCD * nout;
  const int64_t g = b / per_group, r = b % per_group;
  cb = (int) (g * NXCD + r % NXCD);
  io = (int) (r / NXCD);
}

// Column state layout: state[(field * Nmax + k) * ncol + column]; fields
// RHG_CF_* (rhg_colfield), then per continuum its lower-level population
// n_i, then per line chi_l, eta_l, adamp, 1/vbroad, then the depth-scale
// input.  A kernel addresses one (column, depth) point with
// st = state + k * ncol + column and field stride fs = Nmax * ncol.
__host__ __device__ constexpr int cont_field(int ci) { return RHG_CF_NFIXED + ci; }

// index of the first per-line field in the column state
__host__ __device__ __forceinline__ int line_field_base(int ncont) {
  return RHG_CF_NFIXED + ncont;
}
__host__ __device__ __forceinline__ int state_nfield(int ncont, int nline) {
  return RHG_CF_NFIXED + ncont + 4 * nline + 1;
}

// Saha-Boltzmann ratio of level i to level 0 (before normalisation),
// rh/ltepops.c:87-100
__device__ __forceinline__ double level_ratio(const DevModel &M, int l0, int i, double T,
                                              double cNe_T) {
  double dE = ldc(M.level_E)[l0 + i] - ldc(M.level_E)[l0];
  double gi0 = ldc(M.level_g)[l0 + i] / ldc(M.level_g)[l0];
  int dZ = ldc(M.level_stage)[l0 + i] - ldc(M.level_stage)[l0];
  double dE_kT = dE / (KBOLTZMANN * T);
  double r = gi0 * exp(-dE_kT);
  for (int m = 1; m <= dZ; m++) r /= cNe_T;
  return r;
}

// 1 + sum of the level ratios of an atom (the Saha-Boltzmann normalisation)
__device__ __forceinline__ double atom_sum(const DevModel &M, int l0, int nl, double T,
                                           double cNe_T) {
  double sum = 1.0;
  for (int i = 1; i < nl; i++) sum += level_ratio(M, l0, i, T, cNe_T);
  return sum;
}

// The line list of one wavelength as the profile code reads it: slot i
// (0 <= i < n) is a line of the wavelength's list with its Doppler offset
// (lambda/lambda0 - 1) c, and its Zeeman components are c0(i) .. c1(i)-1.
// LineTabG reads the model in global memory; LineTabL (below) reads a copy
// staged in LDS by the kernel.
struct LineTabG {
  static constexpr bool grouped = false;
  const DevModel &M;
  int p0, n;
  // entries p0 .. p0 + n - 1 of the model's wavelength line lists
  __device__ __forceinline__ LineTabG(const DevModel &M_, int p0_, int n_) : M(M_), p0(p0_), n(n_) {}
  __device__ __forceinline__ int line(int i) const { return ldc(M.lam_line_idx)[p0 + i]; }
  __device__ __forceinline__ double dv(int i) const { return ldc(M.lam_line_dv)[p0 + i]; }
  __device__ __forceinline__ bool damped(int i) const { return ldc(M.line_Grad)[line(i)] != 0.0; }
  __device__ __forceinline__ bool polar(int i) const { return ldc(M.line_polar)[line(i)] != 0; }
  __device__ __forceinline__ double larmor(int i) const { return ldc(M.line_larmor)[line(i)]; }
  __device__ __forceinline__ int c0(int i) const { return ldc(M.line_comp_ptr)[line(i)]; }
  __device__ __forceinline__ int c1(int i) const { return ldc(M.line_comp_ptr)[line(i) + 1]; }
  __device__ __forceinline__ double shift(int c) const { return ldc(M.comp_shift)[c]; }
  __device__ __forceinline__ double strength(int c) const { return ldc(M.comp_strength)[c]; }
  __device__ __forceinline__ int q(int c) const { return ldc(M.comp_q)[c]; }
};

// A wavelength's line list staged in LDS (one copy per wavefront; every
// lane reads the same address, so the reads are broadcasts).  Counts and
// indices are made wave-uniform so that the loops over lines and
// components run as scalar loops.  A line's components are stored grouped
// by their Delta M (sigma-, pi, sigma+), each group in the model's order:
// each group feeds its own sum, so the sums see the same terms in the same
// order as from the ungrouped list, and the component loop needs no
// per-component branch.
struct LineRecL {
  double dv, larmor;
  int line, flags, c0, c1;   // flags: 1 damped, 2 polarizable
  int cpi, csp;              // first pi and first sigma+ component
};

struct LineTabL {
  static constexpr bool grouped = true;
  const LineRecL *rec;
  const double *shift_, *strength_;
  const int *q_;
  int n;
  __device__ __forceinline__ int line(int i) const {
    return __builtin_amdgcn_readfirstlane(rec[i].line);
  }
  __device__ __forceinline__ double dv(int i) const { return rec[i].dv; }
  __device__ __forceinline__ bool damped(int i) const {
    return __builtin_amdgcn_readfirstlane(rec[i].flags) & 1;
  }
  __device__ __forceinline__ bool polar(int i) const {
    return __builtin_amdgcn_readfirstlane(rec[i].flags) & 2;
  }
  __device__ __forceinline__ double larmor(int i) const { return rec[i].larmor; }
  __device__ __forceinline__ int c0(int i) const { return __builtin_amdgcn_readfirstlane(rec[i].c0); }
  __device__ __forceinline__ int c1(int i) const { return __builtin_amdgcn_readfirstlane(rec[i].c1); }
  __device__ __forceinline__ int cpi(int i) const { return __builtin_amdgcn_readfirstlane(rec[i].cpi); }
  __device__ __forceinline__ int csp(int i) const { return __builtin_amdgcn_readfirstlane(rec[i].csp); }
  __device__ __forceinline__ double shift(int c) const { return shift_[c]; }
  __device__ __forceinline__ double strength(int c) const { return strength_[c]; }
  __device__ __forceinline__ int q(int c) const { return __builtin_amdgcn_readfirstlane(q_[c]); }
};

// LDS bytes of a staged line list of at most maxl lines and maxc components
__host__ __device__ __forceinline__ size_t linetab_bytes(int maxl, int maxc) {
  return (size_t) maxl * sizeof(LineRecL) + (size_t) maxc * (2 * sizeof(double) + sizeof(int));
}

// Stage wavelength ls's line list into smem (linetab_bytes(maxl, maxc)
// bytes, 8-byte aligned) with the 64 lanes of a one-wavefront workgroup.
// The caller guarantees that the list fits (the host takes maxl and maxc
// as the maxima over the model's wavelengths).  Ends with a barrier.
__device__ __forceinline__ LineTabL stage_lines(const DevModel &M, int ls, char *smem, int maxl,
                                                int maxc) {
  LineRecL *rec = (LineRecL *) smem;
  double *shift = (double *) (smem + (size_t) maxl * sizeof(LineRecL));
  double *strength = shift + maxc;
  int *q = (int *) (strength + maxc);
  const int p0 = ldc(M.lam_line_ptr)[ls], n = ldc(M.lam_line_ptr)[ls + 1] - p0;
  const int lane = threadIdx.x;
  for (int i = lane; i < n; i += 64) {
    const int l = ldc(M.lam_line_idx)[p0 + i];
    const bool damped = ldc(M.line_Grad)[l] != 0.0, polar = ldc(M.line_polar)[l] != 0;
    rec[i].dv = ldc(M.lam_line_dv)[p0 + i];
    rec[i].larmor = ldc(M.line_larmor)[l];
    rec[i].line = l;
    rec[i].flags = (damped ? 1 : 0) | (polar ? 2 : 0);
    // only a damped polarizable line uses its components
    rec[i].c1 = (damped && polar) ? ldc(M.line_comp_ptr)[l + 1] - ldc(M.line_comp_ptr)[l] : 0;
  }
  __syncthreads();
  // offsets: c0(i) = sum of the counts before slot i
  for (int i = lane; i < n; i += 64) {
    int off = 0;
    for (int k = 0; k < i; k++) off += rec[k].c1;
    rec[i].c0 = off;
  }
  __syncthreads();
  for (int i = lane; i < n; i += 64) rec[i].c1 += rec[i].c0;
  __syncthreads();
  // components grouped by Delta M = -1, 0, +1, each in the model's order
  for (int i = 0; i < n; i++) {
    const int a = rec[i].c0, cnt = rec[i].c1 - a;
    const int g0 = ldc(M.line_comp_ptr)[rec[i].line];
    int nsm = 0, npi = 0;
    for (int c = 0; c < cnt; c++) {
      const int qc = ldc(M.comp_q)[g0 + c];
      nsm += qc == -1;
      npi += qc == 0;
    }
    for (int c = lane; c < cnt; c += 64) {
      const int qc = ldc(M.comp_q)[g0 + c];
      // position of component c within its group
      int pos = qc == -1 ? 0 : (qc == 0 ? nsm : nsm + npi);
      for (int b = 0; b < c; b++) pos += ldc(M.comp_q)[g0 + b] == qc;
      shift[a + pos] = ldc(M.comp_shift)[g0 + c];
      strength[a + pos] = ldc(M.comp_strength)[g0 + c];
      q[a + pos] = qc;
    }
    if (lane == 0) {
      rec[i].cpi = a + nsm;
      rec[i].csp = a + nsm + npi;
    }
  }
  __syncthreads();
  LineTabL t;
  t.rec = rec;
  t.shift_ = shift;
  t.strength_ = strength;
  t.q_ = q;
  t.n = __builtin_amdgcn_readfirstlane(n);
  return t;
}

// acc += a b for the Zeeman component sums.  In the FP32 form's
// translation units (FMA contraction allowed) an explicit FMA, so that the
// LDS-table and global-table loops round identically whatever the compiler
// would contract; elsewhere a product and a sum, in the reference's order.
template <typename T>
__device__ __forceinline__ void acc_prod(T &acc, const T &a, const T &b) { acc += a * b; }
__device__ __forceinline__ void acc_prod(double &acc, const double &a, const double &b) {
#ifdef RHG_FP32_FORM_TU
  acc = fma(a, b, acc);
#else
  acc += a * b;
#endif
}

// Profile of slot i of a wavelength's line list (line l) at one depth,
// RLKProfile rh/kurucz.c:729-827 for the emergent ray (to_obs = 1), in the
// number type T (double, float, or a dual number of either).  F(field)
// returns a column-state field at that depth as a T; the Doppler offset
// (lambda/lambda0 - 1) c comes from the line table in FP64.  Returns phi;
// phi_Q,U,V are set for a polarizable line with damping, and with MO (the
// magneto-optical terms) the Faraday-Voigt profiles psi[3] = psi_Q,U,V
// (rh/kurucz.c:811-819).  The Humlicek functions of the Zeeman components
// are evaluated in TV (T, or float in the FP32 form), their sums and
// combinations in T.
template <typename T, typename TV = T, bool MO = false, typename Field, typename Tab>
__device__ __forceinline__ T line_profile_t(const Tab &tab, int i, int lb, const Field &F, T mu,
                                            T vz, int moving, T &phi_Q, T &phi_U, T &phi_V,
                                            T *psi = nullptr) {
  const int l = tab.line(i);
  const T adamp = F(lb + 4 * l + 2), ivb = F(lb + 4 * l + 3);
  T v = T(tab.dv(i)) * ivb;
  if (moving) v += (mu * vz) * ivb;
  const T sv = ivb * T(INV_SQRTPI);
  T phi;
  if (tab.damped(i)) {
    if (tab.polar(i)) {
      const T cosg = F(RHG_CF_COSG);
      const T sin2_gamma = T(1.0) - cosg * cosg;
      const T vB = T(tab.larmor(i)) * F(RHG_CF_B) * ivb;
      T phi_sm = T(0), phi_pi = T(0), phi_sp = T(0), H, Fv;
      T psi_sm = T(0), psi_pi = T(0), psi_sp = T(0);
      const int c0 = tab.c0(i), c1 = tab.c1(i);
      if constexpr (Tab::grouped) {
        // one Delta M group [a, b) into its sums; the next component's
        // shift and strength are read while this one's profile is formed
        auto group = [&](int a, int b, T &phi_g, T &psi_g) {
          double sh = a < b ? tab.shift(a) : 0.0, sg = a < b ? tab.strength(a) : 0.0;
          for (int nz = a; nz < b; nz++) {
            const double shc = sh, sgc = sg;
            if (nz + 1 < b) {
              sh = tab.shift(nz + 1);
              sg = tab.strength(nz + 1);
            }
            TV Fw;
            const T u = v - T(shc) * vB;
            const int reg = humlicek_region(val(adamp), val(u));
            const T Hc = (T) VoigtEval<TV>::eval((TV) adamp, (TV) u, reg, &Fw);
            acc_prod(phi_g, T(sgc), Hc);
            if constexpr (MO) acc_prod(psi_g, T(sgc), (T) Fw);
          }
        };
        group(c0, tab.cpi(i), phi_sm, psi_sm);
        group(tab.cpi(i), tab.csp(i), phi_pi, psi_pi);
        group(tab.csp(i), c1, phi_sp, psi_sp);
      } else {
        for (int nz = c0; nz < c1; nz++) {
          TV Fw;
          const T u = v - T(tab.shift(nz)) * vB;
          const int reg = humlicek_region(val(adamp), val(u));
          H = (T) VoigtEval<TV>::eval((TV) adamp, (TV) u, reg, &Fw);
          Fv = (T) Fw;
          const int q = tab.q(nz);
          const T sg = T(tab.strength(nz));
          acc_prod(q == -1 ? phi_sm : (q == 0 ? phi_pi : phi_sp), sg, H);
          if constexpr (MO) acc_prod(q == -1 ? psi_sm : (q == 0 ? psi_pi : psi_sp), sg, Fv);
        }
      }
      const T phi_sigma = phi_sp + phi_sm;
      const T phi_delta = T(0.5) * phi_pi - T(0.25) * phi_sigma;
      phi = (phi_delta * sin2_gamma + T(0.5) * phi_sigma) * sv;
      phi_Q = phi_delta * sin2_gamma * F(RHG_CF_COS2X) * sv;
      phi_U = phi_delta * sin2_gamma * F(RHG_CF_SIN2X) * sv;
      phi_V = T(0.5) * (phi_sp - phi_sm) * cosg * sv;
      if constexpr (MO) {
        const T psi_sigma = psi_sp + psi_sm;
        const T psi_delta = T(0.5) * psi_pi - T(0.25) * psi_sigma;
        psi[0] = psi_delta * sin2_gamma * F(RHG_CF_COS2X) * sv;
        psi[1] = psi_delta * sin2_gamma * F(RHG_CF_SIN2X) * sv;
        psi[2] = T(0.5) * (psi_sp - psi_sm) * cosg * sv;
      }
    } else {
      phi = voigt_armstrong_t(adamp, v) * sv;
    }
  } else {
    phi = (fabs(v) <= T(MAX_GAUSS_DOPPLER)) ? rexp(-v * v) : T(0);
    phi = phi * sv;
  }
  return phi;
}

// Entry e (line l) of wavelength ls's list, from the model (the census and
// the log gf response functions address entries directly).
template <typename T, typename TV = T, bool MO = false, typename Field>
__device__ __forceinline__ T line_profile_f(const DevModel &M, int e, int l, const Field &F, T mu,
                                            T vz, int moving, T &phi_Q, T &phi_U, T &phi_V,
                                            T *psi = nullptr) {
  (void) l;
  return line_profile_t<T, TV, MO>(LineTabG(M, e, 1), 0, line_field_base(M.ncont), F, mu, vz,
                                   moving, phi_Q, phi_U, phi_V, psi);
}

// Kurucz-line opacity and emissivity of the emergent ray (to_obs = 1) at
// one wavelength and one depth, rh/kurucz.c:600-715: the sum over the
// wavelength's line list (tab) of line_profile_t times the line's chi_l,
// eta_l.  chiI/etaI and chiQ[3]/etaQ[3] accumulate from zero; with MO also
// the magneto-optical chip[3] (chi_l psi_Q,U,V, rh/kurucz.c:711-715).
template <typename T, typename TV = T, bool MO = false, typename Field, typename Tab>
__device__ __forceinline__ void line_point_t(const Tab &tab, int ncont, const Field &F, T mu,
                                             T &chiI, T *chiQ, T &etaI, T *etaQ, int moving,
                                             T *chip = nullptr) {
  const int lb = line_field_base(ncont);
  chiI = T(0); etaI = T(0);
  chiQ[0] = chiQ[1] = chiQ[2] = T(0);
  etaQ[0] = etaQ[1] = etaQ[2] = T(0);
  if constexpr (MO) chip[0] = chip[1] = chip[2] = T(0);
  const T vz = F(RHG_CF_VLOS);
  for (int i = 0; i < tab.n; i++) {
    const int l = tab.line(i);
    const T chi_l = F(lb + 4 * l), eta_l = F(lb + 4 * l + 1);
    T phi_Q = T(0), phi_U = T(0), phi_V = T(0), psi[3] = {T(0), T(0), T(0)};
    const T phi = line_profile_t<T, TV, MO>(tab, i, lb, F, mu, vz, moving, phi_Q, phi_U, phi_V, psi);
    if (phi != T(0)) {
      chiI += chi_l * phi;
      etaI += eta_l * phi;
      if (tab.polar(i) && tab.damped(i)) {
        chiQ[0] += chi_l * phi_Q;
        chiQ[1] += chi_l * phi_U;
        chiQ[2] += chi_l * phi_V;
        etaQ[0] += eta_l * phi_Q;
        etaQ[1] += eta_l * phi_U;
        etaQ[2] += eta_l * phi_V;
        if constexpr (MO) {
          chip[0] += chi_l * psi[0];
          chip[1] += chi_l * psi[1];
          chip[2] += chi_l * psi[2];
        }
      }
    }
  }
}

// the same over wavelength ls's list in the model
template <typename T, typename TV = T, bool MO = false, typename Field>
__device__ __forceinline__ void line_point_f(const DevModel &M, int ls, const Field &F, T mu,
                                             T &chiI, T *chiQ, T &etaI, T *etaQ, int moving,
                                             T *chip = nullptr) {
  const int p0 = ldc(M.lam_line_ptr)[ls];
  line_point_t<T, TV, MO>(LineTabG(M, p0, ldc(M.lam_line_ptr)[ls + 1] - p0), M.ncont, F, mu, chiI, chiQ,
                          etaI, etaQ, moving, chip);
}

// the same with the fields read from the column state at st (stride fs)
template <typename R>
__device__ __forceinline__ void line_point(const DevModel &M, int ls, const double *st, int64_t fs,
                                           R mu, R &chiI, R *chiQ, R &etaI, R *etaQ, int moving) {
  auto F = [st, fs](int f) { return (R) st[f * fs]; };
  line_point_f<R>(M, ls, F, mu, chiI, chiQ, etaI, etaQ, moving);
}

// Work census of one emergent ray (column, wavelength) over depth indices
// 1..N-1: the same line window, Doppler offsets and Voigt branches as
// line_point(), with counts in place of arithmetic (slots: RHG_WORK_*).
__device__ __forceinline__ void line_census(const DevModel &M, int ls, const double *st, int64_t fs,
                                            double mu, int moving, int32_t *cnt) {
  const int p0 = ldc(M.lam_line_ptr)[ls], p1 = ldc(M.lam_line_ptr)[ls + 1];
  const int lb = line_field_base(M.ncont);
  const double vz = st[RHG_CF_VLOS * fs];
  for (int e = p0; e < p1; e++) {
    const int l = ldc(M.lam_line_idx)[e];
    const double adamp = st[(lb + 4 * l + 2) * fs], ivb = st[(lb + 4 * l + 3) * fs];
    double v = ldc(M.lam_line_dv)[e] * ivb;
    if (moving) v += (mu * vz) * ivb;
    cnt[RHG_WORK_LINE]++;
    if (ldc(M.line_Grad)[l] != 0.0) {
      if (ldc(M.line_polar)[l] != 0) {
        const double vB = ldc(M.line_larmor)[l] * st[RHG_CF_B * fs] * ivb;
        for (int nz = ldc(M.line_comp_ptr)[l]; nz < ldc(M.line_comp_ptr)[l + 1]; nz++)
          cnt[RHG_WORK_VOIGT_I + humlicek_region<double>(adamp, v - ldc(M.comp_shift)[nz] * vB)]++;
      } else {
        cnt[RHG_WORK_ARMSTRONG]++;
      }
    } else {
      cnt[RHG_WORK_GAUSS]++;
    }
  }
}

template <int NN>
__global__ void k_state(const DevModel *M, const double *atmos, const int32_t *ndep,
                        const int32_t *perm, int ncol, int Nmax, double mu, int atm_scale,
                        double *state, int nfield);
__global__ void k_geometry(const DevModel *M, const int32_t *ndep, int ncol, int Nmax, int atm_scale,
                           double mu, double *state, int nfield, int32_t *moving_out,
                           double *chiref_out);
__global__ void k_scales(const DevModel *Mp, const int32_t *ndep, int ncol, int Nmax, int atm_scale,
                         const double *state, int nfield, double *tau, double *height,
                         double *cmass);
template <typename R, bool GRID>
__global__ void k_continuum(const DevModel *M, const int32_t *ndep, int ncol, int Nmax,
                            const double *state, int nfield, const int32_t *out_ls, int nout,
                            R *cchi, R *ceta, const int32_t *blk_grid, const int32_t *gtile,
                            const int32_t *gmask, int gwords);
template <typename R, bool MO, bool LT>
__global__ void k_formal_stokes(const DevModel *M, const int32_t *ndep, const int32_t *moving,
                                int ncol, int Nmax, const double *state, const int32_t *out_ls,
                                const int32_t *io_list, int nlist, int nout, const R *cchi,
                                const R *ceta, double mu, double *stokes, const int32_t *perm,
                                const int32_t *blk_grid, const int32_t *gptr, int tab_maxl,
                                int tab_maxc);
template <typename R>
__global__ void k_formal_scalar(const DevModel *M, const int32_t *ndep, const int32_t *moving,
                                int ncol, int Nmax, const double *state, const int32_t *out_ls,
                                const int32_t *io_list, int nlist, int nout, const R *cchi,
                                const R *ceta, double mu, double *stokes, const int32_t *perm,
                                const int32_t *blk_grid, const int32_t *gptr);
__global__ void k_rf_loggf(const DevModel *M, const int32_t *ndep, const int32_t *moving, int ncol,
                           int Nmax, const double *state, const int32_t *out_ls,
                           const int32_t *io_list, int nlist, int nout, const double *cchi,
                           const double *ceta, double mu, int npar, const int32_t *par_line,
                           double *Iin, double *rf);
constexpr int RF_NCARRY = 13;  // carried scalars entering a DELO step (StepCarry)
// row shifts of the perturbed copies of a batch (two per parameter)
struct RfShiftList {
  int32_t row[2 * RHG_RF_NPARAM];
  double h[2 * RHG_RF_NPARAM];
};
__global__ void k_rf_perturb(const double *atmos, double *out, int64_t n, int Nmax,
                             RfShiftList shift);
__global__ void k_rf_rep_ndep(const int32_t *ndep, int ncol, int32_t *out);
template <typename O>
__global__ void k_rf_diff(const double *S, int64_t W, int64_t off_p, int64_t off_m, int ncol,
                          int64_t nrows, double scale, O *out);
template <typename R>
__global__ void k_rf_forward(const DevModel *M, const int32_t *ndep, const int32_t *moving, int ncol,
                             int Nmax, const double *state, const int32_t *out_ls,
                             const int32_t *io_list, int nlist, int nout, const R *cchi,
                             const R *ceta, double mu, double *stokes, double *car);
template <typename R>
__global__ void k_rf_reverse(const DevModel *M, const int32_t *ndep, const int32_t *moving, int ncol,
                             int Nmax, const double *state, const int32_t *out_ls,
                             const int32_t *io_list, int nlist, const R *cchi, const R *ceta,
                             double mu, const double *car, R *aP, R *aZ, R *aB);
// the response-function parameters of one call (rhg_rf_param order)
struct RfParamList {
  int32_t npar;
  int32_t p[RHG_RF_NPARAM];
};
template <typename R, int NP>
__global__ void k_rf_chain(const DevModel *M, const int32_t *ndep, const int32_t *moving, int ncol,
                           int Nmax, const double *state, const double *dstate, const double *chiref,
                           const double *dchiref, const R *cchi, const R *ceta, const R *dcchi,
                           const R *dceta, const int32_t *out_ls, const int32_t *io_list, int nlist,
                           int nout, double mu, int atm_scale, RfParamList pl, const R *aP,
                           const R *aZ, const R *aB, R *rf, int64_t ldr);
__global__ void k_census(const DevModel *M, const int32_t *ndep, const int32_t *moving, int ncol,
                         int Nmax, const double *state, int nfield, const int32_t *out_ls,
                         int nout, double mu, int32_t *census);
__global__ void k_census_sum(const int32_t *census, int ncol, int nout, int64_t *col_sum);
// column reordering for batches of uneven depth counts: records of len
// values per column, gathered into (k_gather_cols) or scattered back from
// (k_scatter_cols) the sorted order; k_scatter_last for arrays whose last
// index is the column
__global__ void k_gather_cols(const double *src, double *dst, const int32_t *perm, int ncol,
                              int64_t len);
__global__ void k_scatter_cols(const double *src, double *dst, const int32_t *perm, int ncol,
                               int64_t len);
template <typename T>
__global__ void k_scatter_last(const T *src, T *dst, const int32_t *perm, int ncol, int64_t nouter);

}  // namespace rhg
