// NLTE populations of one active atom per column: the accelerated lambda
// iteration of rhf1d() (rh/iterate.c:48-143) as a sequence of batched
// kernels over all columns at once.
//
// Set-up (once per batch, after k_state of the background model)
//   k_nlte_pre      (column, depth): scattering opacity at the reference
//                   wavelength (the depth scale), LTE populations, Doppler
//                   widths, damping and the collision matrix.
//   k_nlte_bg       (column, depth): scattering opacity at every wavelength,
//                   added to the LTE continuum opacity.
//   k_nlte_profile  (column, depth, line): absorption profiles for every
//                   ray and direction, and their normalisation.
// Iteration (per pass, over the columns still iterating)
//   k_nlte_bezier   (column, wavelength with a line): the rays of the
//   k_nlte_feautrier (column, continuum-only wavelength)
//                   wavelength with the diagonal operator; new mean
//                   intensity and, with the operator, the radiative rate
//                   terms of every transition active there.
//   k_nlte_stateq   (column, depth): rate matrix = collisions + the rate
//                   terms summed over the transitions' wavelengths; the
//                   statistical-equilibrium solve.
//   k_nlte_ng       (column): Ng acceleration and the maximum relative
//                   change.
#include "rhg_nlte.hpp"

namespace rhg {

namespace {

// rh/planck.c:38-63 for a wavelength in nm
__device__ __forceinline__ double planck_lam(double T, double lambda) {
  const double hc_kla = (HPLANCK * CLIGHT) / (KBOLTZMANN * NM_TO_M * lambda);
  const double twohnu3_c2 = (2.0 * HPLANCK * CLIGHT) / ((NM_TO_M * lambda) * (NM_TO_M * lambda) *
                                                         (NM_TO_M * lambda));
  const double hc_Tkla = hc_kla / T;
  return (hc_Tkla <= MAX_EXPONENT) ? twohnu3_c2 / (exp(hc_Tkla) - 1.0) : 0.0;
}

// Fritsch & Butland centred derivative with the reference's divisions,
// rh/bezier_aux.c:34-51
// a / b as a times the hardware reciprocal estimate refined by two Newton
// steps (relative error ~1e-16, not correctly rounded): the quotients of
// the scalar Bezier rays, where the IEEE division sequence is a long
// dependent chain per use
__device__ __forceinline__ double ndiv(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  return a * fma(r, e, r);
}

__device__ __forceinline__ double cent_div(double dsup, double dsdn, double up, double c,
                                           double dn) {
  const double fim1 = ndiv(c - up, dsup), fi = ndiv(dn - c, dsdn);
  if (fim1 * fi > 0) {
    const double alpha = 0.333333333333333333333333 * (1.0 + ndiv(dsdn, dsdn + dsup));
    return ndiv(fim1 * fi, (1.0 - alpha) * fim1 + alpha * fi);
  }
  return 0.0;
}

__device__ __forceinline__ double max0(double x) { return x > 0.0 ? x : 0.0; }

// RH's splineEval with precomputed second derivatives, rh/spline.c:70-97
__device__ double spline_eval_M(int n, const double *x, const double *y, const double *Mv,
                                double xv) {
  if (xv <= x[0]) return y[0];
  if (xv >= x[n - 1]) return y[n - 1];
  const int j = locate_asc(n, x, xv);
  const double hj = x[j + 1] - x[j];
  const double fx = (xv - x[j]) / hj;
  const double fx1 = 1 - fx;
  return fx1 * y[j] + fx * y[j + 1] +
         (fx1 * (fx1 * fx1 - 1) * Mv[j] + fx * (fx * fx - 1) * Mv[j + 1]) * (hj * hj) / 6.0;
}

constexpr double HC = HPLANCK * CLIGHT;
constexpr double FOURPI = 4.0 * PI;
constexpr double HC_4PI = HC / FOURPI;

}  // namespace

// ---------------------------------------------------------------------------
// Set-up per (column, depth).  Scattering opacity at the reference
// wavelength joins chi_ref before convertScales (rh/background.c:461-463);
// LTE populations of the active atom (rh/ltepops.c:33-113), its total and
// Doppler width (rh/readatom.c:213-220); damping of each line
// (rh/broad.c:273-316); collision rates (rh/collision.c:450-760).
__global__ void k_nlte_pre(const DevModel *__restrict__ Mp, const DevNlte *__restrict__ Ap,
                           NlteBufs B) {
  const DevModel &M = *Mp;
  const DevNlte &A = *Ap;
  const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  const int ncol = B.ncol, Nmax = B.Nmax;
  if (p >= (int64_t) ncol * Nmax) return;
  const int ic = (int) (p % ncol), k = (int) (p / ncol);
  if (k >= B.ndep[ic]) return;
  const int64_t fs = (int64_t) Nmax * ncol;
  double *st = B.state + (int64_t) k * ncol + ic;
  const double T = st[RHG_CF_T * fs], ne = st[RHG_CF_NE * fs], nHtot = st[RHG_CF_NHTOT * fs];
  const double nH0 = st[RHG_CF_NH0 * fs], np = st[RHG_CF_NP * fs], nH2 = st[RHG_CF_NH2 * fs];
  const double vturb = st[RHG_CF_VTURB * fs];
  const double c1 = (HPLANCK / (2.0 * PI * M_ELECTRON)) * (HPLANCK / KBOLTZMANN);
  const double cNe_T = 0.5 * ne * pow(c1 / T, 1.5);
  const int64_t pk = (int64_t) k * ncol + ic;

  double nHe = 0.0;
  if (A.he_atom >= 0) {
    const int l0 = ldc(M.atom_level_ptr)[A.he_atom], nl = ldc(M.atom_level_ptr)[A.he_atom + 1] - l0;
    nHe = (ldc(M.atom_abund)[A.he_atom] * nHtot) / atom_sum(M, l0, nl, T, cNe_T);
  }
  B.nHe[pk] = nHe;
  {
    const int lr = M.ref_index;
    double sca = ne * A.sigma_thomson;
    sca += A.ray_H[lr] * nH0;
    if (A.he_atom >= 0) sca += A.ray_He[lr] * nHe;
    sca += A.ray_H2[lr] * nH2;
    sca *= ldc(M.lamconst)[(int64_t) lr * RHG_LC_NCONST + RHG_LC_SCATT_FUDGE];
    st[RHG_CF_HEIGHT * fs] += sca;
  }

  const int nl = A.nl, a = A.atom_index, l0 = ldc(M.atom_level_ptr)[a];
  const double ntotal = ldc(M.atom_abund)[a] * nHtot;
  const double n0 = ntotal / atom_sum(M, l0, nl, T, cNe_T);
  for (int i = 0; i < nl; i++) {
    const double ns = (i == 0) ? n0 : level_ratio(M, l0, i, T, cNe_T) * n0;
    B.nstar[(int64_t) i * fs + pk] = ns;
    B.pops[(int64_t) i * fs + pk] = ns;
  }
  B.ntot[pk] = ntotal;
  const double vtherm = 2.0 * KBOLTZMANN / (AMU * A.weight);
  const double vbroad = sqrt(vtherm * T + vturb * vturb);
  B.vbroad[pk] = vbroad;

  for (int l = 0; l < A.nline; l++) {
    const double cDop = (NM_TO_M * A.ln_lambda0[l]) / (4.0 * PI);
    double Qelast = 0.0;
    if (A.ln_vdw[l]) {
      // rh/broad.c:87-100 (Unsold), 125-136 (Barklem), 145
      double G = A.ln_vdw[l] == 2
                     ? A.ln_vdw_cross[l] * pow(T, (1.0 - A.ln_vdw_alpha[l]) / 2.0) +
                           A.ln_vdw_cross_he[l] * pow(T, 0.3)
                     : A.ln_vdw_cross[l] * pow(T, 0.3);
      G *= nH0;
      Qelast += G;
    }
    if (A.ln_stark[l] == 1) {
      Qelast += A.ln_stark_c[l] * ne;
    } else if (A.ln_stark[l] == 2) {
      const double vrel = pow(A.ln_stark_C[l] * T, 0.16666667) * A.ln_stark_Cm[l];
      Qelast += A.ln_stark_c[l] * vrel * ne;
    }
    B.adamp[(int64_t) l * fs + pk] = (A.ln_Grad[l] + Qelast) * cDop / vbroad;
  }

  for (int q = 0; q < nl * nl; q++) B.Cmat[(int64_t) q * fs + pk] = 0.0;
  const double C0 = ((E_RYDBERG / sqrt(M_ELECTRON)) * PI * (RBOHR * RBOHR)) *
                    sqrt(8.0 / (PI * KBOLTZMANN));
  for (int r = 0; r < A.ncoll; r++) {
    const int i1 = A.coll_i[r], j1 = A.coll_j[r], c0 = A.coll_ptr[r], nT = A.coll_ptr[r + 1] - c0;
    const double Cv = nT > 2 ? spline_eval_M(nT, A.coll_T + c0, A.coll_c + c0, A.coll_M + c0, T)
                             : linear1(nT, A.coll_T + c0, A.coll_c + c0, T);
    double *Cij = B.Cmat + (int64_t) (i1 * nl + j1) * fs + pk;
    double *Cji = B.Cmat + (int64_t) (j1 * nl + i1) * fs + pk;
    const double nsi = B.nstar[(int64_t) i1 * fs + pk], nsj = B.nstar[(int64_t) j1 * fs + pk];
    double Cdown, Cup;
    switch (A.coll_type[r]) {
      case 0:   // OMEGA
        Cdown = C0 * ne * Cv / (A.g[j1] * sqrt(T));
        *Cij += Cdown;
        *Cji += Cdown * nsj / nsi;
        break;
      case 1:   // CE
        Cdown = Cv * ne * (A.g[i1] / A.g[j1]) * sqrt(T);
        *Cij += Cdown;
        *Cji += Cdown * nsj / nsi;
        break;
      case 2:   // CI
        Cup = Cv * ne * exp(-(A.E[j1] - A.E[i1]) / (KBOLTZMANN * T)) * sqrt(T);
        *Cji += Cup;
        *Cij += Cup * nsi / nsj;
        break;
      case 3:   // CP
        Cdown = np * Cv;
        *Cij += Cdown;
        *Cji += Cdown * nsj / nsi;
        break;
      case 4:   // CH
        Cup = nH0 * Cv;
        *Cji += Cup;
        *Cij += Cup * nsi / nsj;
        break;
      case 5:   // CH0
        *Cij += nH0 * Cv;
        break;
      case 6:   // CH+
        *Cji += np * Cv;
        break;
    }
  }
}

// Scattering opacity (Thomson and Rayleigh on H, He and H2) at every output
// wavelength: stored, and added to the LTE continuum opacity
// (rh/background.c:342-465).
__global__ void k_nlte_bg(const DevModel *__restrict__ Mp, const DevNlte *__restrict__ Ap,
                          NlteBufs B, const int32_t *__restrict__ out_ls) {
  const DevNlte &A = *Ap;
  const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  const int ncol = B.ncol, Nmax = B.Nmax;
  if (p >= (int64_t) ncol * Nmax) return;
  const int ic = (int) (p % ncol), k = (int) (p / ncol);
  if (k >= B.ndep[ic]) return;
  const int64_t fs = (int64_t) Nmax * ncol;
  const double *st = B.state + (int64_t) k * ncol + ic;
  const double ne = st[RHG_CF_NE * fs], nH0 = st[RHG_CF_NH0 * fs], nH2 = st[RHG_CF_NH2 * fs];
  const double nHe = B.nHe[(int64_t) k * ncol + ic];
  const DevModel &M = *Mp;
  for (int io = 0; io < B.nout; io++) {
    const int ls = out_ls[io];
    double sca = ne * A.sigma_thomson;
    sca += A.ray_H[ls] * nH0;
    if (A.he_atom >= 0) sca += A.ray_He[ls] * nHe;
    sca += A.ray_H2[ls] * nH2;
    // the scattering fudge factor, rh/background.c:453-462
    sca *= ldc(M.lamconst)[(int64_t) ls * RHG_LC_NCONST + RHG_LC_SCATT_FUDGE];
    const int64_t q = (int64_t) io * fs + (int64_t) k * ncol + ic;
    B.csca[q] = sca;
    B.cchi[q] += sca;
  }
}

// Field-free absorption profiles of the atom's lines for every ray and
// direction and their normalisation over the line's quadrature,
// rh/profile.c:67-371 (Voigt function of Armstrong, rh/voigt.c:126-243).
__global__ void k_nlte_profile(const DevModel *__restrict__ Mp, const DevNlte *__restrict__ Ap,
                               NlteBufs B) {
  const DevModel &M = *Mp;
  const DevNlte &A = *Ap;
  const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  const int ncol = B.ncol, Nmax = B.Nmax;
  if (p >= (int64_t) ncol * Nmax * A.nline) return;
  const int ic = (int) (p % ncol);
  const int64_t q = p / ncol;
  const int k = (int) (q % Nmax), l = (int) (q / Nmax);
  if (k >= B.ndep[ic]) return;
  const int64_t fs = (int64_t) Nmax * ncol, pk = (int64_t) k * ncol + ic;
  const double vb = B.vbroad[pk], vz = B.state[RHG_CF_VLOS * fs + pk];
  const double adamp = B.adamp[(int64_t) l * fs + pk];
  const int t = A.ln_tr[l], nrays = A.nrays;
  const double lambda0 = A.ln_lambda0[l];
  double wphi = 0.0;
  for (int la = 0; la < A.tr_nlam[t]; la++) {
    const int ls = A.tr_nblue[t] + la;
    const double lamla = ldc(M.lamconst)[(int64_t) ls * RHG_LC_NCONST + RHG_LC_LAMBDA];
    const double wl = A.tr_wlam[A.tr_ptr[t] + la];
    const double v = (lamla - lambda0 - 0.0) * CLIGHT / (vb * lambda0);
    for (int mu = 0; mu < nrays; mu++) {
      const double wlamu = wl * 0.5 * A.wmu[mu];
      const double vlos = A.muz[mu] * vz / vb;
      for (int to_obs = 0; to_obs <= 1; to_obs++) {
        const double sign = to_obs ? 1.0 : -1.0;
        const double vk = v + sign * vlos;
        const double ph = voigt_armstrong(adamp, vk) * 1.0 / (SQRTPI * vb);
        B.phi[((((int64_t) (A.ln_phi[l] + la) * nrays + mu) * 2 + to_obs) * Nmax) * ncol + pk] = ph;
        wphi += ph * wlamu;
      }
    }
  }
  B.wphi[(int64_t) l * fs + pk] = 1.0 / wphi;
}

// ---------------------------------------------------------------------------
// Radiative rate terms at one depth for the rays of one direction
// (addtoCoupling and addtoGamma, rh/fillgamma.c:82-332), summed over the
// rays.  The members of the active set are described per wavelength by
// SetInfo; per depth the kernel passes the members' populations, gij and
// quadrature weights, and per ray the absorption V of each member.
template <int MS>
struct SetInfo {
  int n;                                  // members
  int ti[MS], tj[MS], ent[MS], line[MS], prow[MS];
  double tw[MS], gline[MS], cV[MS], alpha[MS], wlam[MS];
  // per member: members with the same lower level / upper level / whose
  // upper level is this member's lower level (with 2hnu^3/c^2 != 0 for the
  // sums, rh/fillgamma.c:251-332), and the count of the last without that
  // condition (rh/fillgamma.c:204-214)
  uint32_t m_up[MS], m_jj[MS], m_ji[MS];
  int ncross[MS];
};

// per-depth member data
template <int MS>
struct SetDepth {
  double ni[MS], nj[MS], gij[MS], wla[MS];
};

// accumulate the rate terms of ray r (intensity I, operator Psi / chi in
// PsiC, direction weight wmu, absorption V of each member) into sIJ, sJI
template <int MS>
__device__ __forceinline__ void rate_terms(const SetInfo<MS> &si, const SetDepth<MS> &sd,
                                           const double (&V)[MS], double I, double PsiC,
                                           double wmu, double (&sIJ)[MS], double (&sJI)[MS]) {
  double eta = 0.0;
#pragma unroll
  for (int n = 0; n < MS; n++)
    if (n < si.n) eta += si.tw[n] * sd.gij[n] * V[n] * sd.nj[n];
  const double Ieff = I - PsiC * eta;
  double chicc[MS], U[MS];
#pragma unroll
  for (int n = 0; n < MS; n++) {
    chicc[n] = V[n] * sd.wla[n] * (sd.ni[n] - sd.gij[n] * sd.nj[n]);
    U[n] = si.tw[n] * sd.gij[n] * V[n];
  }
#pragma unroll
  for (int n = 0; n < MS; n++) {
    if (n < si.n) {
      double chi_up = 0.0, Uj = 0.0, chi_dn = 0.0, Ui = 0.0;
#pragma unroll
      for (int m = 0; m < MS; m++) {
        if ((si.m_up[n] >> m) & 1u) chi_up += chicc[m];
        if ((si.m_jj[n] >> m) & 1u) { Uj += U[m]; chi_dn += chicc[m]; }
        if ((si.m_ji[n] >> m) & 1u) Ui += U[m];
      }
      const double wlamu = V[n] * sd.wla[n] * wmu;
      double gji = sJI[n], gij = sIJ[n];
      gji += Ieff * wlamu;
      gij += (si.tw[n] + Ieff) * sd.gij[n] * wlamu;
      gij -= chi_up * PsiC * Uj * wmu;
      for (int c = 0; c < si.ncross[n]; c++) gji += chi_dn * PsiC * Ui * wmu;
      sJI[n] = gji;
      sIJ[n] = gij;
    }
  }
}

// The active set of wavelength ls: members, their constants and coupling
// masks.
template <int MS>
__device__ __forceinline__ void set_info(const DevNlte &A, int ls, double lam, SetInfo<MS> &si) {
  const int p0 = A.as_ptr[ls];
  si.n = A.as_ptr[ls + 1] - p0;
  const double twohc = 2.0 * HC / (NM_TO_M * NM_TO_M * NM_TO_M);
#pragma unroll
  for (int n = 0; n < MS; n++) {
    si.ti[n] = si.tj[n] = si.ent[n] = si.line[n] = si.prow[n] = 0;
    si.tw[n] = si.gline[n] = si.cV[n] = si.alpha[n] = si.wlam[n] = 0.0;
    if (n < si.n) {
      const int t = A.as_tr[p0 + n], la = ls - A.tr_nblue[t];
      si.ti[n] = A.tr_i[t];
      si.tj[n] = A.tr_j[t];
      si.ent[n] = A.tr_ptr[t] + la;
      si.wlam[n] = A.tr_wlam[A.tr_ptr[t] + la];
      if (A.tr_type[t] == 0) {
        const int l = A.tr_line[t];
        si.line[n] = l;
        si.prow[n] = (A.ln_phi[l] + la) * A.nrays;
        si.gline[n] = A.ln_Bji[l] / A.ln_Bij[l];
        si.cV[n] = HC_4PI * A.ln_Bij[l] * 1.0;
        si.tw[n] = A.ln_Aji[l] / A.ln_Bji[l];
      } else {
        si.line[n] = -1;
        si.tw[n] = twohc / (lam * lam * lam);
        si.alpha[n] = A.tr_alpha[A.tr_ptr[t] + la];
      }
    }
  }
#pragma unroll
  for (int n = 0; n < MS; n++) {
    uint32_t up = 0, jj = 0, ji = 0;
    int nc = 0;
#pragma unroll
    for (int m = 0; m < MS; m++) {
      if (m < si.n) {
        if (si.tw[m] != 0.0) {
          if (si.ti[m] == si.ti[n]) up |= 1u << m;
          if (si.tj[m] == si.tj[n]) jj |= 1u << m;
          if (si.tj[m] == si.ti[n]) ji |= 1u << m;
        }
        if (si.tj[m] == si.ti[n]) nc++;
      }
    }
    si.m_up[n] = up;
    si.m_jj[n] = jj;
    si.m_ji[n] = ji;
    si.ncross[n] = nc;
  }
}

// ---------------------------------------------------------------------------
// Wavelengths with an active line, one (column, wavelength) per thread:
// the rays with the scalar cubic Bezier method and its diagonal operator
// (rh/rhf1d/bezier_1D.c:306-541), both directions, the new mean intensity
// and with eval_op the radiative rate terms (rh/rhf1d/formal.c:44-346,
// rh/fillgamma.c:82-332).
//
// The RB rays of a direction are swept together, depth by depth: RB
// independent recursions per thread.  The opacity and source function of
// each ray are evaluated two depths ahead of the recursion into a window
// in registers, so no per-depth array is kept.  At each depth the rays'
// contributions to J and to each member's rate terms are summed in
// registers and stored once per direction: the inward sweep writes them,
// the outward sweep adds its own.  J is double-buffered per column
// (jpar): the previous J enters the source function, the new one is built.
//
// Workgroups map to (64 columns of the active list, list wavelength) by
// xcd_tile, so a block of columns' state and populations stay in one
// XCD's L2 across the wavelengths.
// With active sets of at most two members and RB = 3 rays per sweep the
// kernel is held to two waves per SIMD (256 VGPRs, 22 spilled): measured
// on a solve of 4096 columns of the Ca II atom, 10% faster than RB = 2 at
// three waves (fewer sweeps re-read the background and update the rate
// sums) and than RB = 3 or 5 at one.
template <int RB, int MS>
__global__ void __launch_bounds__(64, (MS <= 2 ? 2 : 1))
k_nlte_bezier(const DevModel *__restrict__ Mp, const DevNlte *__restrict__ Ap, NlteBufs B,
              const int32_t *__restrict__ act, int nact, const int32_t *__restrict__ out_ls,
              const int32_t *__restrict__ io_list, int nlist, int il_off, int eval_op) {
  const DevModel &M = *Mp;
  const DevNlte &A = *Ap;
  const int ncb = (nact + 63) / 64;
  int cb, il;
  xcd_tile(blockIdx.x, nlist, cb, il);
  if (cb >= ncb) return;
  const int slot = cb * 64 + threadIdx.x;
  if (slot >= nact) return;
  const int ic = act[slot], io = io_list[il], ls = out_ls[io];
  const int ncol = B.ncol, Nmax = B.Nmax, N = B.ndep[ic], nrays = A.nrays;
  const int64_t fs = (int64_t) Nmax * ncol;
#define STF(f, k) B.state[(int64_t) (f) * fs + (int64_t) (k) * ncol + ic]
#define PK(k) ((int64_t) (k) * ncol + ic)
  const int64_t wo = (int64_t) io * fs, jb = (int64_t) B.nout * fs;
  const int par = B.jpar[ic];
  const double *Jold = B.J + par * jb + wo;
  double *Jnew = B.J + (1 - par) * jb + wo;
  const double *chic = B.cchi + wo, *etac = B.ceta + wo, *scac = B.csca + wo;
  const double lam = ldc(M.lamconst)[(int64_t) ls * RHG_LC_NCONST + RHG_LC_LAMBDA];
  const double hc_k = HC / (KBOLTZMANN * NM_TO_M);
  SetInfo<MS> si;
  set_info<MS>(A, ls, lam, si);
  const double B2_0 = planck_lam(STF(RHG_CF_T, N - 2), lam);
  const double B2_1 = planck_lam(STF(RHG_CF_T, N - 1), lam);
  double dJmax = 0.0;

  // member data at depth k
  auto depth = [&](int k, SetDepth<MS> &sd) {
    const double T = STF(RHG_CF_T, k);
#pragma unroll
    for (int n = 0; n < MS; n++) {
      sd.ni[n] = sd.nj[n] = sd.gij[n] = sd.wla[n] = 0.0;
      if (n < si.n) {
        sd.ni[n] = B.pops[(int64_t) si.ti[n] * fs + PK(k)];
        sd.nj[n] = B.pops[(int64_t) si.tj[n] * fs + PK(k)];
        if (si.line[n] >= 0) {
          sd.gij[n] = si.gline[n];
          sd.wla[n] = si.wlam[n] * B.wphi[(int64_t) si.line[n] * fs + PK(k)] / HC_4PI;
        } else {
          sd.gij[n] = B.nstar[(int64_t) si.ti[n] * fs + PK(k)] /
                      B.nstar[(int64_t) si.tj[n] * fs + PK(k)] * exp(-hc_k / (lam * T));
          sd.wla[n] = FOURPI / HPLANCK * (si.wlam[n] / lam);
        }
      }
    }
  };
  // absorption per population of each member for ray mu, direction dir
  auto absorb = [&](int k, int mu, int dir, double (&V)[MS]) {
#pragma unroll
    for (int n = 0; n < MS; n++) {
      V[n] = 0.0;
      if (n < si.n) {
        if (si.line[n] >= 0)
          V[n] = si.cV[n] *
                 B.phi[(((int64_t) (si.prow[n] + mu) * 2 + dir) * Nmax) * ncol + PK(k)];
        else
          V[n] = si.alpha[n];
      }
    }
  };

  for (int c0 = 0; c0 < nrays; c0 += RB) {
    const int nrc = nrays - c0 < RB ? nrays - c0 : RB;
    const bool last_chunk = c0 + RB >= nrays;
    double zmu[RB], wmu[RB];
#pragma unroll
    for (int r = 0; r < RB; r++) {
      zmu[r] = 1.0;
      wmu[r] = 0.0;
      if (r < nrc) {
        zmu[r] = 1.0 / A.muz[c0 + r];
        wmu[r] = 0.5 * A.wmu[c0 + r];
      }
    }
    for (int dir = 0; dir <= 1; dir++) {
      const bool first = (c0 == 0 && dir == 0);
      // opacity and source function of the rays at depth k
      auto opac = [&](int k, double (&C)[RB], double (&S)[RB]) {
        SetDepth<MS> sd;
        depth(k, sd);
        const double cc = chic[PK(k)], ee = etac[PK(k)], ss = scac[PK(k)], jd = Jold[PK(k)];
#pragma unroll
        for (int r = 0; r < RB; r++) {
          C[r] = 1.0;
          S[r] = 0.0;
          if (r < nrc) {
            double V[MS];
            absorb(k, c0 + r, dir, V);
            double chi = 0.0, eta = 0.0;
#pragma unroll
            for (int n = 0; n < MS; n++) {
              if (n < si.n) {
                chi += V[n] * (sd.ni[n] - sd.gij[n] * sd.nj[n]);
                eta += si.tw[n] * sd.gij[n] * V[n] * sd.nj[n];
              }
            }
            C[r] = chi + cc;
            double s = eta + ee;
            s += ss * jd;
            S[r] = ndiv(s, C[r]);
          }
        }
      };
      // J and rate terms of depth k
      auto emit = [&](int k, const double (&I)[RB], const double (&Psi)[RB],
                      const double (&C)[RB]) {
        double js = 0.0;
#pragma unroll
        for (int r = 0; r < RB; r++)
          if (r < nrc) js += wmu[r] * I[r];
        double *jp = Jnew + PK(k);
        const double jn = first ? js : *jp + js;
        *jp = jn;
        if (last_chunk && dir == 1) {
          const double dJ = fabs(1.0 - Jold[PK(k)] / jn);
          dJmax = dJmax > dJ ? dJmax : dJ;
        }
        if (!eval_op) return;
        SetDepth<MS> sd;
        depth(k, sd);
        double sIJ[MS], sJI[MS];
#pragma unroll
        for (int n = 0; n < MS; n++) sIJ[n] = sJI[n] = 0.0;
#pragma unroll
        for (int r = 0; r < RB; r++) {
          if (r < nrc) {
            double V[MS];
            absorb(k, c0 + r, dir, V);
            rate_terms<MS>(si, sd, V, I[r], ndiv(Psi[r], C[r]), wmu[r], sIJ, sJI);
          }
        }
#pragma unroll
        for (int n = 0; n < MS; n++) {
          if (n < si.n) {
            double *aij = B.acc + ((int64_t) si.ent[n] * 2) * fs + PK(k);
            double *aji = aij + fs;
            *aij = first ? sIJ[n] : *aij + sIJ[n];
            *aji = first ? sJI[n] : *aji + sJI[n];
          }
        }
      };
      // |z[a] - z[b]| for neighbouring depths
      auto dzb = [&](int a, int b) { return STF(RHG_CF_DZ, a > b ? a : b); };
      const int dk = dir ? -1 : 1, ks = dir ? N - 1 : 0, ke = dir ? 0 : N - 1;
      double Cm[RB], C0[RB], Cp[RB], Cn[RB], Sm[RB], S0[RB], Sp[RB], Sn[RB];
      double I_upw[RB], dchi_c[RB], dtau_uw[RB], dS_up[RB], I[RB], Psi[RB];
      opac(ks, Cm, Sm);
      opac(ks + dk, C0, S0);
      opac(ks + 2 * dk, Cp, Sp);
      {
        const double dz0 = dzb(ks, ks + dk);
#pragma unroll
        for (int r = 0; r < RB; r++) {
          const double dt = 0.5 * zmu[r] * (Cm[r] + C0[r]) * dz0;
          I_upw[r] = dir ? B2_1 - (B2_0 - B2_1) / dt : 0.0;
          Psi[r] = 0.0;
        }
        emit(ks, I_upw, Psi, Cm);
        const int k = ks + dk;
        const double dzu = dzb(k, k - dk), dzd = dzb(k + dk, k);
#pragma unroll
        for (int r = 0; r < RB; r++) {
          const double dsup = dzu * zmu[r], dsdn = dzd * zmu[r];
          const double dchi_up = ndiv(C0[r] - Cm[r], dsup);
          dchi_c[r] = cent_div(dsup, dsdn, Cm[r], C0[r], Cp[r]);
          const double c1 = max0(C0[r] - (dsup * (1.0 / 3.0)) * dchi_c[r]);
          const double c2 = max0(Cm[r] + (dsup * (1.0 / 3.0)) * dchi_up);
          dtau_uw[r] = dsup * (C0[r] + Cm[r] + c1 + c2) * 0.25;
          dS_up[r] = ndiv(S0[r] - Sm[r], dtau_uw[r]);
        }
      }
      for (int k = ks + dk; k != ke + dk; k += dk) {
        if (k != ke) {
          const double dzd = dzb(k + dk, k);
          const bool far = abs(k - ke) > 1;
          double dzd2 = 0.0;
          if (far) {
            opac(k + 2 * dk, Cn, Sn);
            dzd2 = dzb(k + 2 * dk, k + dk);
          } else {
#pragma unroll
            for (int r = 0; r < RB; r++) {
              Cn[r] = 1.0;
              Sn[r] = 0.0;
            }
          }
#pragma unroll
          for (int r = 0; r < RB; r++) {
            const double dsdn = dzd * zmu[r];
            const double dchi_dn = far ? cent_div(dsdn, dzd2 * zmu[r], C0[r], Cp[r], Cn[r])
                                       : ndiv(Cp[r] - C0[r], dsdn);
            double c1 = max0(C0[r] + (dsdn * (1.0 / 3.0)) * dchi_c[r]);
            double c2 = max0(Cp[r] - (dsdn * (1.0 / 3.0)) * dchi_dn);
            const double dtau_dw = dsdn * (C0[r] + Cp[r] + c1 + c2) * 0.25;
            const double dt03 = dtau_uw[r] * (1.0 / 3.0);
            double alpha, beta, gamma, theta, eps;
            bezier3_coeffs(dtau_uw[r], alpha, beta, gamma, theta, eps);
            const double dS_c = cent_div(dtau_uw[r], dtau_dw, Sm[r], S0[r], Sp[r]);
            c1 = max0(S0[r] - dt03 * dS_c);
            c2 = max0(Sm[r] + dt03 * dS_up[r]);
            I[r] = I_upw[r] * eps + alpha * S0[r] + beta * Sm[r] + gamma * c1 + theta * c2;
            Psi[r] = alpha + gamma;
            I_upw[r] = I[r];
            dchi_c[r] = dchi_dn;
            dtau_uw[r] = dtau_dw;
            dS_up[r] = dS_c;
          }
        } else {
          const double dzu = dzb(k, k - dk);
#pragma unroll
          for (int r = 0; r < RB; r++) {
            const double dt = 0.5 * zmu[r] * (C0[r] + Cm[r]) * dzu;
            const double dS_uw = -(S0[r] - Sm[r]) / dt;
            double w[3];
            w3(dt, w);
            I[r] = (1.0 - w[0]) * I_upw[r] + w[0] * S0[r] + w[1] * dS_uw;
            Psi[r] = w[0] - w[1] / dt;
          }
        }
        emit(k, I, Psi, C0);
#pragma unroll
        for (int r = 0; r < RB; r++) {
          Cm[r] = C0[r];
          C0[r] = Cp[r];
          Cp[r] = Cn[r];
          Sm[r] = S0[r];
          S0[r] = Sp[r];
          Sp[r] = Sn[r];
        }
      }
    }
  }
  B.dJ[(int64_t) (il_off + il) * ncol + ic] = dJmax;
#undef STF
#undef PK
}

#define RHG_INST_BEZ(MS)                                                                        \
  template __global__ void k_nlte_bezier<NLTE_RB, MS>(const DevModel *, const DevNlte *, NlteBufs, \
                                                      const int32_t *, int, const int32_t *,     \
                                                      const int32_t *, int, int, int);
RHG_INST_BEZ(2)
RHG_INST_BEZ(4)
RHG_INST_BEZ(8)
#undef RHG_INST_BEZ

// ---------------------------------------------------------------------------
// Wavelengths with continua only, one (column, wavelength) per thread: one
// opacity for all rays, and per ray the Feautrier solution with its
// two-sided diagonal operator (rh/rhf1d/feautrier.c:56-202; P is the mean
// of the two directions, weighted by the full ray weight).  The opacity,
// source function and the elimination coefficients F and z of each ray
// live in the thread's scratch rows; the back substitution runs the rays of
// a chunk together and adds the rate terms of each depth as it reaches it.
// MS is the largest active set among the wavelengths of io_list (the member
// loops are unrolled to it).
template <int MS>
__global__ void __launch_bounds__(64)
k_nlte_feautrier(const DevModel *__restrict__ Mp, const DevNlte *__restrict__ Ap, NlteBufs B,
                 const int32_t *__restrict__ act, int nact, const int32_t *__restrict__ out_ls,
                 const int32_t *__restrict__ io_list, int nlist, int il_off, int eval_op) {
  const DevModel &M = *Mp;
  const DevNlte &A = *Ap;
  const int ncb = (nact + 63) / 64;
  int cb, il;
  xcd_tile(blockIdx.x, nlist, cb, il);
  if (cb >= ncb) return;
  const int slot = cb * 64 + threadIdx.x;
  if (slot >= nact) return;
  const int ic = act[slot], io = io_list[il], ls = out_ls[io];
  const int ncol = B.ncol, Nmax = B.Nmax, N = B.ndep[ic], nrays = A.nrays;
  const int64_t fs = (int64_t) Nmax * ncol;
  const int64_t nthr = (int64_t) nlist * ncb * 64;
  double *scr = B.scr + (int64_t) il * ncb * 64 + slot;
#define SCR(s, k) scr[((int64_t) (s) * Nmax + (k)) * nthr]
#define STF(f, k) B.state[(int64_t) (f) * fs + (int64_t) (k) * ncol + ic]
#define PK(k) ((int64_t) (k) * ncol + ic)
  const int64_t wo = (int64_t) io * fs, jb = (int64_t) B.nout * fs;
  const int par = B.jpar[ic];
  const double *Jold = B.J + par * jb + wo;
  double *Jnew = B.J + (1 - par) * jb + wo;
  const double *chic = B.cchi + wo, *etac = B.ceta + wo, *scac = B.csca + wo;
  const double lam = ldc(M.lamconst)[(int64_t) ls * RHG_LC_NCONST + RHG_LC_LAMBDA];
  const double hc_k = HC / (KBOLTZMANN * NM_TO_M);
  SetInfo<MS> si;
  set_info<MS>(A, ls, lam, si);

  auto depth = [&](int k, SetDepth<MS> &sd, double (&V)[MS]) {
    const double T = STF(RHG_CF_T, k);
#pragma unroll
    for (int n = 0; n < MS; n++) {
      sd.ni[n] = sd.nj[n] = sd.gij[n] = sd.wla[n] = V[n] = 0.0;
      if (n < si.n) {
        sd.ni[n] = B.pops[(int64_t) si.ti[n] * fs + PK(k)];
        sd.nj[n] = B.pops[(int64_t) si.tj[n] * fs + PK(k)];
        if (si.line[n] >= 0) {
          // a line at a Feautrier wavelength (not produced by the set-up):
          // its profile of the first ray
          sd.gij[n] = si.gline[n];
          sd.wla[n] = si.wlam[n] * B.wphi[(int64_t) si.line[n] * fs + PK(k)] / HC_4PI;
          V[n] = si.cV[n] * B.phi[(((int64_t) si.prow[n] * 2) * Nmax) * ncol + PK(k)];
        } else {
          sd.gij[n] = B.nstar[(int64_t) si.ti[n] * fs + PK(k)] /
                      B.nstar[(int64_t) si.tj[n] * fs + PK(k)] * exp(-hc_k / (lam * T));
          sd.wla[n] = FOURPI / HPLANCK * (si.wlam[n] / lam);
          V[n] = si.alpha[n];
        }
      }
    }
  };

  for (int k = 0; k < N; k++) {
    SetDepth<MS> sd;
    double V[MS];
    depth(k, sd, V);
    double chi = 0.0, eta = 0.0;
#pragma unroll
    for (int n = 0; n < MS; n++) {
      if (n < si.n) {
        chi += V[n] * (sd.ni[n] - sd.gij[n] * sd.nj[n]);
        eta += si.tw[n] * sd.gij[n] * V[n] * sd.nj[n];
      }
    }
    const double chi_t = chi + chic[PK(k)];
    double S = eta + etac[PK(k)];
    S += scac[PK(k)] * Jold[PK(k)];
    SCR(NS_CHI, k) = chi_t;
    SCR(NS_S, k) = S / chi_t;
  }
  const double B2_0 = planck_lam(STF(RHG_CF_T, N - 2), lam);
  const double B2_1 = planck_lam(STF(RHG_CF_T, N - 1), lam);
#define CHI(k) SCR(NS_CHI, k)
#define SRC(k) SCR(NS_S, k)
#define HGT(k) STF(RHG_CF_HEIGHT, k)
  // The rays in chunks of up to RF: each ray's forward elimination into its
  // own F and z rows, then the back substitution of the chunk's rays
  // together, so that each depth forms the member data once and updates J
  // and the rate sums once per chunk.  The rays are summed in ray order, as
  // one ray at a time would sum them.
  constexpr int RF = NLTE_FE_RAYS;
  for (int c0 = 0; c0 < nrays; c0 += RF) {
    const int nrc = nrays - c0 < RF ? nrays - c0 : RF;
    double wmu[RF], zmu[RF];
#pragma unroll
    for (int r = 0; r < RF; r++) {
      wmu[r] = r < nrc ? A.wmu[c0 + r] : 0.0;
      zmu[r] = r < nrc ? 0.5 / A.muz[c0 + r] : 1.0;
    }
    auto emit = [&](int k, const double (&P)[RF], const double (&Psi)[RF]) {
      double *jp = Jnew + PK(k);
      double j = c0 == 0 ? 0.0 : *jp;
#pragma unroll
      for (int r = 0; r < RF; r++)
        if (r < nrc) j = (c0 + r == 0 ? 0.0 : j) + P[r] * wmu[r];
      *jp = j;
      if (!eval_op) return;
      SetDepth<MS> sd;
      double V[MS], aIJ[MS], aJI[MS];
      depth(k, sd, V);
#pragma unroll
      for (int n = 0; n < MS; n++) {
        aIJ[n] = aJI[n] = 0.0;
        if (n < si.n && c0 > 0) {
          const double *aij = B.acc + ((int64_t) si.ent[n] * 2) * fs + PK(k);
          aIJ[n] = aij[0];
          aJI[n] = aij[fs];
        }
      }
#pragma unroll
      for (int r = 0; r < RF; r++) {
        if (r < nrc) {
          double sIJ[MS], sJI[MS];
#pragma unroll
          for (int n = 0; n < MS; n++) sIJ[n] = sJI[n] = 0.0;
          rate_terms<MS>(si, sd, V, P[r], Psi[r] / CHI(k), wmu[r], sIJ, sJI);
#pragma unroll
          for (int n = 0; n < MS; n++) {
            aIJ[n] = (c0 + r == 0 ? 0.0 : aIJ[n]) + sIJ[n];
            aJI[n] = (c0 + r == 0 ? 0.0 : aJI[n]) + sJI[n];
          }
        }
      }
#pragma unroll
      for (int n = 0; n < MS; n++) {
        if (n < si.n) {
          double *aij = B.acc + ((int64_t) si.ent[n] * 2) * fs + PK(k);
          aij[0] = aIJ[n];
          aij[fs] = aJI[n];
        }
      }
    };
    auto dtau = [&](int r, int k) {
      return zmu[r] * (CHI(k) + CHI(k + 1)) * (HGT(k) - HGT(k + 1));
    };
    // forward elimination, rh/rhf1d/feautrier.c:56-202 (dark top,
    // thermalised bottom, STANDARD order), ray r into rows NS_F + 2 r and
    // NS_Z + 2 r
    double F[RF], z[RF], dtm1[RF], abc0[RF], C1_0[RF];
#pragma unroll
    for (int r = 0; r < RF; r++) {
      if (r < nrc) {
        const double dt0 = dtau(r, 0);
        abc0[r] = 1.0 + 2.0 * 1.0 / dt0;
        C1_0[r] = 2.0 / (dt0 * dt0);
        const double S0 = SRC(0) + 2.0 * 0.0 / ((1.0 + 0.0) * dt0);
        F[r] = abc0[r] / C1_0[r];
        z[r] = S0 / (abc0[r] + C1_0[r]);
        SCR(NS_F + 2 * r, 0) = F[r];
        SCR(NS_Z + 2 * r, 0) = z[r];
        dtm1[r] = dt0;
      }
    }
    for (int k = 1; k < N - 1; k++) {
#pragma unroll
      for (int r = 0; r < RF; r++) {
        if (r < nrc) {
          const double dtk = dtau(r, k);
          const double dtau_mid = 0.5 * (dtk + dtm1[r]);
          const double A1 = 1.0 / (dtau_mid * dtm1[r]), C1 = 1.0 / (dtau_mid * dtk);
          F[r] = (1.0 + A1 * F[r] / (1.0 + F[r])) / C1;
          z[r] = (SRC(k) + A1 * z[r]) / (C1 * (1.0 + F[r]));
          SCR(NS_F + 2 * r, k) = F[r];
          SCR(NS_Z + 2 * r, k) = z[r];
          dtm1[r] = dtk;
        }
      }
    }
    // back substitution with the two-sided operator
    double P[RF], Psi[RF], G[RF], dtk[RF];
#pragma unroll
    for (int r = 0; r < RF; r++) {
      P[r] = Psi[r] = G[r] = 0.0;
      dtk[r] = 1.0;
      if (r < nrc) {
        const double dtN = dtm1[r];   // dtau[N-2]
        const double hN = B2_1 - (B2_0 - B2_1) / dtN;
        const double abcN = 1.0 + 2.0 * 1.0 / dtN;
        const double A1N = 2.0 / (dtN * dtN);
        const double SN = SRC(N - 1) + 2.0 * hN / ((1.0 + 0.0) * dtN);
        const double FN2 = SCR(NS_F + 2 * r, N - 2);
        P[r] = (SN + A1N * SCR(NS_Z + 2 * r, N - 2)) / (abcN + A1N * (FN2 / (1.0 + FN2)));
        Psi[r] = 1.0 / (abcN + A1N * FN2 / (1.0 + FN2));
        G[r] = abcN / A1N;   // G[N-1]
        dtk[r] = dtN;        // dtau[k] for k = N-2
      }
    }
    emit(N - 1, P, Psi);
    for (int k = N - 2; k >= 1; k--) {
#pragma unroll
      for (int r = 0; r < RF; r++) {
        if (r < nrc) {
          const double Fk = SCR(NS_F + 2 * r, k), Fm = SCR(NS_F + 2 * r, k - 1);
          P[r] = P[r] / (1.0 + Fk) + SCR(NS_Z + 2 * r, k);
          const double dtkm = dtau(r, k - 1);
          const double dtau_mid = 0.5 * (dtk[r] + dtkm);
          const double A1 = 1.0 / (dtau_mid * dtkm), C1 = 1.0 / (dtau_mid * dtk[r]);
          Psi[r] = 1.0 / (1.0 + A1 * Fm / (1.0 + Fm) + C1 * G[r] / (1.0 + G[r]));
          G[r] = (1.0 + C1 * G[r] / (1.0 + G[r])) / A1;
          dtk[r] = dtkm;
        }
      }
      emit(k, P, Psi);
    }
#pragma unroll
    for (int r = 0; r < RF; r++) {
      if (r < nrc) {
        P[r] = P[r] / (1.0 + SCR(NS_F + 2 * r, 0)) + SCR(NS_Z + 2 * r, 0);
        Psi[r] = 1.0 / (abc0[r] + C1_0[r] * G[r] / (1.0 + G[r]));
      }
    }
    emit(0, P, Psi);
  }
  double dJmax = 0.0;
  for (int k = 0; k < N; k++) {
    const double dJ = fabs(1.0 - Jold[PK(k)] / Jnew[PK(k)]);
    dJmax = dJmax > dJ ? dJmax : dJ;
  }
  B.dJ[(int64_t) (il_off + il) * ncol + ic] = dJmax;
#undef HGT
#undef CHI
#undef SRC
#undef SCR
#undef STF
#undef PK
}
#define RHG_INST_FE(MS)                                                                          \
  template __global__ void k_nlte_feautrier<MS>(const DevModel *, const DevNlte *, NlteBufs,     \
                                                const int32_t *, int, const int32_t *,           \
                                                const int32_t *, int, int, int);
RHG_INST_FE(2)
RHG_INST_FE(4)
RHG_INST_FE(6)
RHG_INST_FE(8)
#undef RHG_INST_FE

// ---------------------------------------------------------------------------
// The emergent spectrum after convergence, pyrh's final ray
// (rh/rhf1d/pyrh_solveray.c:82-107): one ray at the caller's mu with the
// converged populations, profiles for that ray (Armstrong, as the set-up's),
// background scattering from the last J, no update of J.  One (column,
// output wavelength) per thread: the opacity and source function into the
// scratch rows, then the outward Bezier ray (wavelengths with a line) or
// the Feautrier solution at mu (rh/rhf1d/formal.c:44-346).  I goes to
// stokes[(column * 4) * nout + io]; Q, U, V are zero (field-free).
__global__ void __launch_bounds__(64)
k_nlte_emergent(const DevModel *__restrict__ Mp, const DevNlte *__restrict__ Ap, NlteBufs B,
                const int32_t *__restrict__ out_ls, const int32_t *__restrict__ io_list, int nlist,
                double mu, double *__restrict__ stokes) {
  constexpr int MS = NLTE_MAXSET;
  const DevModel &M = *Mp;
  const DevNlte &A = *Ap;
  const int ncol = B.ncol, Nmax = B.Nmax, ncb = (ncol + 63) / 64;
  int cb, il;
  xcd_tile(blockIdx.x, nlist, cb, il);
  if (cb >= ncb) return;
  const int ic = cb * 64 + threadIdx.x;
  if (ic >= ncol) return;
  const int io = io_list[il], ls = out_ls[io], N = B.ndep[ic];
  const int64_t fs = (int64_t) Nmax * ncol;
  const int64_t nthr = (int64_t) nlist * ncb * 64;
  double *scr = B.scr + (int64_t) il * ncb * 64 + ic;
#define SCR(s, k) scr[((int64_t) (s) * Nmax + (k)) * nthr]
#define STF(f, k) B.state[(int64_t) (f) * fs + (int64_t) (k) * ncol + ic]
#define PK(k) ((int64_t) (k) * ncol + ic)
  const int64_t wo = (int64_t) io * fs, jb = (int64_t) B.nout * fs;
  const double *Jl = B.J + B.jpar[ic] * jb + wo;
  const double *chic = B.cchi + wo, *etac = B.ceta + wo, *scac = B.csca + wo;
  const double lam = ldc(M.lamconst)[(int64_t) ls * RHG_LC_NCONST + RHG_LC_LAMBDA];
  const double hc_k = HC / (KBOLTZMANN * NM_TO_M);
  SetInfo<MS> si;
  set_info<MS>(A, ls, lam, si);
  const bool bez = si.n > 0 && A.as_bezier[ls];
  for (int k = 0; k < N; k++) {
    const double T = STF(RHG_CF_T, k), vb = B.vbroad[PK(k)], vz = STF(RHG_CF_VLOS, k);
    double chi = 0.0, eta = 0.0;
#pragma unroll
    for (int n = 0; n < MS; n++) {
      if (n < si.n) {
        const double ni = B.pops[(int64_t) si.ti[n] * fs + PK(k)];
        const double nj = B.pops[(int64_t) si.tj[n] * fs + PK(k)];
        double V, gij;
        if (si.line[n] >= 0) {
          const int l = si.line[n];
          const double lambda0 = A.ln_lambda0[l];
          const double v = (lam - lambda0 - 0.0) * CLIGHT / (vb * lambda0);
          const double vlos = mu * vz / vb;
          const double ph = voigt_armstrong(B.adamp[(int64_t) l * fs + PK(k)], v + 1.0 * vlos) *
                            1.0 / (SQRTPI * vb);
          gij = si.gline[n];
          V = si.cV[n] * ph;
        } else {
          gij = B.nstar[(int64_t) si.ti[n] * fs + PK(k)] / B.nstar[(int64_t) si.tj[n] * fs + PK(k)] *
                exp(-hc_k / (lam * T));
          V = si.alpha[n];
        }
        chi += V * (ni - gij * nj);
        eta += si.tw[n] * gij * V * nj;
      }
    }
    const double chi_t = chi + chic[PK(k)];
    double S = eta + etac[PK(k)];
    S += scac[PK(k)] * Jl[PK(k)];
    SCR(NS_CHI, k) = chi_t;
    SCR(NS_S, k) = S / chi_t;
  }
  const double B2_0 = planck_lam(STF(RHG_CF_T, N - 2), lam);
  const double B2_1 = planck_lam(STF(RHG_CF_T, N - 1), lam);
#define CHI(k) SCR(NS_CHI, k)
#define SRC(k) SCR(NS_S, k)
  double Iout;
  if (bez) {
    // outward ray from the thermalised bottom, rh/rhf1d/bezier_1D.c:306-541
    const double zmu = 1.0 / mu;
    auto dzb = [&](int k) { return STF(RHG_CF_DZ, k); };   // |z[k-1] - z[k]|
    double dtau_uw = 0.5 * zmu * (CHI(N - 1) + CHI(N - 2)) * dzb(N - 1);
    double I_upw = B2_1 - (B2_0 - B2_1) / dtau_uw;
    int k = N - 2;
    double dsup = dzb(k + 1) * zmu, dsdn = dzb(k) * zmu;
    double dchi_up = (CHI(k) - CHI(k + 1)) / dsup;
    double dchi_c = cent_div(dsup, dsdn, CHI(k + 1), CHI(k), CHI(k - 1));
    double c1 = max0(CHI(k) - (dsup * (1.0 / 3.0)) * dchi_c);
    double c2 = max0(CHI(k + 1) + (dsup * (1.0 / 3.0)) * dchi_up);
    dtau_uw = dsup * (CHI(k) + CHI(k + 1) + c1 + c2) * 0.25;
    double dS_up = (SRC(k) - SRC(k + 1)) / dtau_uw;
    for (k = N - 2; k >= 1; k--) {
      dsdn = dzb(k) * zmu;
      double dchi_dn;
      if (k > 1) {
        const double dsdn2 = dzb(k - 1) * zmu;
        dchi_dn = cent_div(dsdn, dsdn2, CHI(k), CHI(k - 1), CHI(k - 2));
      } else {
        dchi_dn = (CHI(k - 1) - CHI(k)) / dsdn;
      }
      c1 = max0(CHI(k) + (dsdn * (1.0 / 3.0)) * dchi_c);
      c2 = max0(CHI(k - 1) - (dsdn * (1.0 / 3.0)) * dchi_dn);
      const double dtau_dw = dsdn * (CHI(k) + CHI(k - 1) + c1 + c2) * 0.25;
      const double dt03 = dtau_uw / 3.0;
      double alpha, beta, gamma, theta, eps;
      bezier3_coeffs(dtau_uw, alpha, beta, gamma, theta, eps);
      const double dS_c = cent_div(dtau_uw, dtau_dw, SRC(k + 1), SRC(k), SRC(k - 1));
      c1 = max0(SRC(k) - dt03 * dS_c);
      c2 = max0(SRC(k + 1) + dt03 * dS_up);
      I_upw = I_upw * eps + alpha * SRC(k) + beta * SRC(k + 1) + gamma * c1 + theta * c2;
      dchi_c = dchi_dn;
      dtau_uw = dtau_dw;
      dS_up = dS_c;
    }
    {
      const double dt = 0.5 * zmu * (CHI(0) + CHI(1)) * dzb(1);
      const double dS_uw = -(SRC(0) - SRC(1)) / dt;
      double w[3];
      w3(dt, w);
      Iout = (1.0 - w[0]) * I_upw + w[0] * SRC(0) + w[1] * dS_uw;
    }
    (void) dchi_up;
  } else {
    // Feautrier at mu, rh/rhf1d/feautrier.c:56-202; the emergent intensity
    // is (1 + f0) P[0] - h0 / (1 + r0) with r0 = h0 = 0
#define HGT(k) STF(RHG_CF_HEIGHT, k)
    const double zmu = 0.5 / mu;
    auto dtau = [&](int k) { return zmu * (CHI(k) + CHI(k + 1)) * (HGT(k) - HGT(k + 1)); };
    const double dt0 = dtau(0);
    const double abc0 = 1.0 + 2.0 * 1.0 / dt0;
    const double C1_0 = 2.0 / (dt0 * dt0);
    const double S0 = SRC(0) + 2.0 * 0.0 / ((1.0 + 0.0) * dt0);
    double F = abc0 / C1_0, z = S0 / (abc0 + C1_0);
    SCR(NS_F, 0) = F;
    SCR(NS_Z, 0) = z;
    double dtm1 = dt0;
    for (int k = 1; k < N - 1; k++) {
      const double dtk = dtau(k);
      const double dtau_mid = 0.5 * (dtk + dtm1);
      const double A1 = 1.0 / (dtau_mid * dtm1), C1 = 1.0 / (dtau_mid * dtk);
      F = (1.0 + A1 * F / (1.0 + F)) / C1;
      z = (SRC(k) + A1 * z) / (C1 * (1.0 + F));
      SCR(NS_F, k) = F;
      SCR(NS_Z, k) = z;
      dtm1 = dtk;
    }
    const double dtN = dtm1;
    const double hN = B2_1 - (B2_0 - B2_1) / dtN;
    const double abcN = 1.0 + 2.0 * 1.0 / dtN;
    const double A1N = 2.0 / (dtN * dtN);
    const double SN = SRC(N - 1) + 2.0 * hN / ((1.0 + 0.0) * dtN);
    const double FN2 = SCR(NS_F, N - 2);
    double P = (SN + A1N * SCR(NS_Z, N - 2)) / (abcN + A1N * (FN2 / (1.0 + FN2)));
    for (int k = N - 2; k >= 0; k--) P = P / (1.0 + SCR(NS_F, k)) + SCR(NS_Z, k);
    Iout = (1.0 + 1.0) * P - 0.0 / (1.0 + 0.0);
#undef HGT
  }
  double *o = stokes + (int64_t) ic * 4 * B.nout + io;
  o[0] = Iout;
  o[B.nout] = 0.0;
  o[2 * (int64_t) B.nout] = 0.0;
  o[3 * (int64_t) B.nout] = 0.0;
#undef CHI
#undef SRC
#undef SCR
#undef STF
#undef PK
}

// the new J becomes the previous one for the columns just solved
__global__ void k_nlte_flip(NlteBufs B, const int32_t *__restrict__ act, int nact) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nact) return;
  const int ic = act[slot];
  B.jpar[ic] = 1 - B.jpar[ic];
}

// max over the list wavelengths of the per-wavelength dJ, per column
__global__ void k_nlte_djmax(NlteBufs B, const int32_t *__restrict__ act, int nact, int nlist,
                             double *__restrict__ djmax) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nact) return;
  const int ic = act[slot];
  double d = 0.0;
  for (int il = 0; il < nlist; il++) {
    const double x = B.dJ[(int64_t) il * B.ncol + ic];
    d = d > x ? d : x;
  }
  djmax[ic] = d;
}

// ---------------------------------------------------------------------------
// Crout LU decomposition with implicit pivoting, back substitution and one
// step of iterative improvement, rh/ludcmp.c:36-177, on an NLP x NLP system
// held in registers (compile-time indices; the dynamic pivot row is
// selected by comparison).  A system of n < NLP equations is padded with
// identity rows and zero right-hand sides, which leaves every operation on
// the first n rows and columns unchanged.
template <int NLP>
__device__ __forceinline__ void lu_decomp_r(double (&a)[NLP][NLP], int (&index)[NLP]) {
  double vv[NLP];
#pragma unroll
  for (int i = 0; i < NLP; i++) {
    double big = 0.0;
#pragma unroll
    for (int j = 0; j < NLP; j++) {
      const double temp = fabs(a[i][j]);
      if (temp > big) big = temp;
    }
    vv[i] = 1.0 / big;
  }
#pragma unroll
  for (int j = 0; j < NLP; j++) {
#pragma unroll
    for (int i = 0; i < j; i++) {
      double sum = a[i][j];
#pragma unroll
      for (int k = 0; k < i; k++) sum -= a[i][k] * a[k][j];
      a[i][j] = sum;
    }
    double big = 0.0;
    int imax = 0;
#pragma unroll
    for (int i = j; i < NLP; i++) {
      double sum = a[i][j];
#pragma unroll
      for (int k = 0; k < j; k++) sum -= a[i][k] * a[k][j];
      a[i][j] = sum;
      const double dum = vv[i] * fabs(sum);
      if (dum >= big) {
        big = dum;
        imax = i;
      }
    }
#pragma unroll
    for (int i = j + 1; i < NLP; i++) {
      if (i == imax) {
#pragma unroll
        for (int k = 0; k < NLP; k++) {
          const double d = a[i][k];
          a[i][k] = a[j][k];
          a[j][k] = d;
        }
        vv[i] = vv[j];
      }
    }
    index[j] = imax;
    if (a[j][j] == 0.0) a[j][j] = 1.0e-20;
    const double dum = 1.0 / a[j][j];
#pragma unroll
    for (int i = j + 1; i < NLP; i++) a[i][j] *= dum;
  }
}

template <int NLP>
__device__ __forceinline__ void lu_backsubst_r(const double (&a)[NLP][NLP],
                                               const int (&index)[NLP], double (&b)[NLP]) {
  int ii = -1;
#pragma unroll
  for (int i = 0; i < NLP; i++) {
    const int ip = index[i];
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < NLP; q++)
      if (q == ip) sum = b[q];
#pragma unroll
    for (int q = 0; q < NLP; q++)
      if (q == ip) b[q] = b[i];
    if (ii >= 0) {
#pragma unroll
      for (int j = 0; j < i; j++)
        if (j >= ii) sum -= a[i][j] * b[j];
    } else if (sum) {
      ii = i;
    }
    b[i] = sum;
  }
#pragma unroll
  for (int i = NLP - 1; i >= 0; i--) {
    double sum = b[i];
#pragma unroll
    for (int j = i + 1; j < NLP; j++) sum -= a[i][j] * b[j];
    b[i] = sum / a[i][i];
  }
}

template <int NLP>
__device__ __forceinline__ void solve_improve_r(double (&a)[NLP][NLP], double (&b)[NLP]) {
  double ac[NLP][NLP], bc[NLP], res[NLP];
  int index[NLP];
#pragma unroll
  for (int i = 0; i < NLP; i++) {
    bc[i] = b[i];
#pragma unroll
    for (int j = 0; j < NLP; j++) ac[i][j] = a[i][j];
  }
  lu_decomp_r<NLP>(a, index);
  lu_backsubst_r<NLP>(a, index, b);
#pragma unroll
  for (int i = 0; i < NLP; i++) {
    res[i] = bc[i];
#pragma unroll
    for (int j = 0; j < NLP; j++) res[i] -= ac[i][j] * b[j];
  }
  lu_backsubst_r<NLP>(a, index, res);
#pragma unroll
  for (int i = 0; i < NLP; i++) b[i] += res[i];
}

// Rate matrix and statistical equilibrium at one (column, depth),
// rh/statequil.c:36-216: Gamma = C + the rate terms of each transition
// summed over its wavelengths in order; the diagonal is minus the column
// sums; the equation of the most populated level (isum = -1, else level
// isum) is replaced by particle conservation.
template <int NLP>
__global__ void __launch_bounds__(64)
k_nlte_stateq(const DevNlte *__restrict__ Ap, NlteBufs B, const int32_t *__restrict__ act,
              int nact, int isum) {
  const DevNlte &A = *Ap;
  const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  const int Nmax = B.Nmax, ncol = B.ncol, nl = A.nl;
  if (p >= (int64_t) nact * Nmax) return;
  const int slot = (int) (p % nact), k = (int) (p / nact);
  const int ic = act[slot];
  if (k >= B.ndep[ic]) return;
  const int64_t fs = (int64_t) Nmax * ncol, pk = (int64_t) k * ncol + ic;
  for (int q = 0; q < nl * nl; q++) B.Gam[(int64_t) q * fs + pk] = B.Cmat[(int64_t) q * fs + pk];
  for (int t = 0; t < A.ntrans; t++) {
    const int i = A.tr_i[t], j = A.tr_j[t], e0 = A.tr_ptr[t], ne = A.tr_nlam[t];
    double *gij = B.Gam + (int64_t) (i * nl + j) * fs + pk;
    double *gji = B.Gam + (int64_t) (j * nl + i) * fs + pk;
    double sij = *gij, sji = *gji;
    for (int e = e0; e < e0 + ne; e++) {
      sij += B.acc[((int64_t) e * 2) * fs + pk];
      sji += B.acc[((int64_t) e * 2 + 1) * fs + pk];
    }
    *gij = sij;
    *gji = sji;
  }
  double a[NLP][NLP], nk[NLP];
#pragma unroll
  for (int i = 0; i < NLP; i++) {
    nk[i] = (i < nl) ? B.pops[(int64_t) i * fs + pk] : 0.0;
#pragma unroll
    for (int j = 0; j < NLP; j++)
      a[i][j] = (i < nl && j < nl) ? B.Gam[(int64_t) (i * nl + j) * fs + pk] : (i == j ? 1.0 : 0.0);
  }
  int ie = 0;
  if (isum == -1) {
    double nmax = 0.0;
#pragma unroll
    for (int i = 0; i < NLP; i++)
      if (i < nl && nk[i] > nmax) {
        nmax = nk[i];
        ie = i;
      }
  } else {
    ie = isum;
  }
#pragma unroll
  for (int i = 0; i < NLP; i++) {
    if (i < nl) {
      a[i][i] = 0.0;
      nk[i] = 0.0;
      double GamDiag = 0.0;
#pragma unroll
      for (int j = 0; j < NLP; j++)
        if (j < nl) GamDiag += a[j][i];
      a[i][i] = -GamDiag;
    }
  }
  const double ntotal = B.ntot[pk];
#pragma unroll
  for (int i = 0; i < NLP; i++) {
    if (i == ie) {
      nk[i] = ntotal;
#pragma unroll
      for (int j = 0; j < NLP; j++)
        if (j < nl) a[i][j] = 1.0;
    }
  }
  solve_improve_r<NLP>(a, nk);
#pragma unroll
  for (int i = 0; i < NLP; i++)
    if (i < nl) B.pops[(int64_t) i * fs + pk] = nk[i];
}

template __global__ void k_nlte_stateq<4>(const DevNlte *, NlteBufs, const int32_t *, int, int);
template __global__ void k_nlte_stateq<6>(const DevNlte *, NlteBufs, const int32_t *, int, int);
template __global__ void k_nlte_stateq<8>(const DevNlte *, NlteBufs, const int32_t *, int, int);
template __global__ void k_nlte_stateq<16>(const DevNlte *, NlteBufs, const int32_t *, int, int);

// ---------------------------------------------------------------------------
// Ng acceleration of one column's populations and the maximum relative
// change, rh/accelerate.c:34-146 and rh/maxchange.c:32-50, over the
// populations flattened level by level (index i * N + k).  The history
// holds norder + 2 solutions; ng_init stored the starting solution in slot
// 0, so before iteration niter (1-based) the count is niter.
__global__ void k_nlte_ng(const DevNlte *__restrict__ Ap, NlteBufs B,
                          const int32_t *__restrict__ act, int nact, int niter, int norder,
                          int ndelay, int nperiod, double *__restrict__ dpops) {
  const DevNlte &A = *Ap;
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nact) return;
  const int ic = act[slot], N = B.ndep[ic], nl = A.nl, Nmax = B.Nmax, ncol = B.ncol;
  const int NS = norder + 2;
  const int64_t fs = (int64_t) Nmax * ncol, hs = (int64_t) nl * fs;
  auto H = [&](int s, int i, int k) -> double & {
    return B.hist[(int64_t) s * hs + (int64_t) i * fs + (int64_t) k * ncol + ic];
  };
  auto POP = [&](int i, int k) -> double & { return B.pops[(int64_t) i * fs + (int64_t) k * ncol + ic]; };
  const int cur = niter % NS;
  for (int i = 0; i < nl; i++)
    for (int k = 0; k < N; k++) H(cur, i, k) = POP(i, k);
  const int count = niter + 1;
  if (norder > 0 && count >= ndelay && !((count - ndelay) % nperiod)) {
    double Am[NLTE_MAXNG][NLTE_MAXNG], bv[NLTE_MAXNG];
#pragma unroll
    for (int i = 0; i < NLTE_MAXNG; i++) {
      bv[i] = 0.0;
#pragma unroll
      for (int j = 0; j < NLTE_MAXNG; j++) Am[i][j] = (i < norder || j < norder) ? 0.0 : (i == j ? 1.0 : 0.0);
    }
    for (int i = 0; i < nl; i++)
      for (int k = 0; k < N; k++) {
        double D[NLTE_MAXNG + 1];
#pragma unroll
        for (int q = 0; q <= NLTE_MAXNG; q++) {
          D[q] = 0.0;
          // count >= ndelay >= norder + 2, so both slots are >= 0
          if (q <= norder) D[q] = H((count - 1 - q) % NS, i, k) - H((count - 2 - q) % NS, i, k);
        }
        const double w = 1.0 / fabs(POP(i, k));
#pragma unroll
        for (int j = 0; j < NLTE_MAXNG; j++) {
          if (j < norder) {
            bv[j] += w * D[0] * (D[0] - D[j + 1]);
#pragma unroll
            for (int ii = 0; ii < NLTE_MAXNG; ii++)
              if (ii < norder) Am[ii][j] += w * (D[j + 1] - D[0]) * (D[ii + 1] - D[0]);
          }
        }
      }
    solve_improve_r<NLTE_MAXNG>(Am, bv);
    const int i0 = (count - 1) % NS;
    for (int i = 0; i < nl; i++)
      for (int k = 0; k < N; k++) {
        double s = POP(i, k);
        const double h0 = H(i0, i, k);
#pragma unroll
        for (int q = 0; q < NLTE_MAXNG; q++)
          if (q < norder) s += bv[q] * (H((count - 2 - q) % NS, i, k) - h0);
        POP(i, k) = s;
        H(i0, i, k) = s;
      }
  }
  double dmax = 0.0;
  {
    const int so = (count - 2) % NS, sc = (count - 1) % NS;
    for (int i = 0; i < nl; i++)
      for (int k = 0; k < N; k++) {
        const double c = H(sc, i, k);
        if (c) {
          const double d = fabs((c - H(so, i, k)) / c);
          dmax = dmax > d ? dmax : d;
        }
      }
  }
  dpops[ic] = dmax;
}

// populations (src: current or LTE) to the caller's layout
// [ncol][nl][Nmax], zero past ndep
__global__ void k_nlte_out(const DevNlte *__restrict__ Ap, NlteBufs B, const double *__restrict__ src,
                           double *__restrict__ out) {
  const DevNlte &A = *Ap;
  const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  const int ncol = B.ncol, Nmax = B.Nmax, nl = A.nl;
  if (p >= (int64_t) ncol * nl * Nmax) return;
  const int k = (int) (p % Nmax);
  const int64_t q = p / Nmax;
  const int i = (int) (q % nl), ic = (int) (q / nl);
  out[p] = k < B.ndep[ic] ? src[((int64_t) i * Nmax + k) * ncol + ic] : 0.0;
}

}  // namespace rhg
