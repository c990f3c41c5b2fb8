// Hydrostatic equilibrium for batches of temperature columns: pyrh's hse()
// (pyrh.pyx:427-487, rh/rhf1d/pyrh_hse.c:67-399) with one lane group per
// column.  Per layer, from the top down, the reference repeats until n_Htot
// settles: rho from n_Htot, n_e from LTE ionisation of every element
// (get_ne, rh/rhf1d/pyrh_background.c:459-549), LTE populations and
// chemical equilibrium, the continuum opacity at the evaluation wavelength
// (pyrh_Background, pyrh_background.c:144-457), p_g from the layer above,
// and n_Htot = (p_g/kT - n_e)/totalAbund.  The layers are a serial chain, so
// the parallelism is over columns; the element sum of the Newton step is
// split over the lanes of a group to give small batches more lanes.
#include "rhg_equil.hpp"
#include "rhg_hse.hpp"

namespace rhg {

namespace {

constexpr int HSE_N_MAX_ELECTRON_ITER = 10;  // rh/rhf1d/pyrh_background.h:1-2
constexpr double HSE_MAX_ELECTRON_ERROR = 1.0E-2;

// What the layer's temperature fixes in the Saha factors: pow(C1/T, 1.5),
// exp(E_ION_HMIN/kT) (these two in this order, as chem_eq takes them),
// pow(C1/T, -1.5) and exp(Uk0 + chi_H/kT).
struct HseT {
  double p15, eHmin, pm15, eH;
};

__device__ __forceinline__ HseT hse_tcache(const DevHse &H, double T, double Uk0) {
  const double C1 = (HPLANCK / (2.0 * PI * M_ELECTRON)) * (HPLANCK / KBOLTZMANN);
  const double kT = KBOLTZMANN * T;
  HseT t;
  t.p15 = pow(C1 / T, 1.5);
  t.eHmin = exp(0.754 * EV / kT);
  t.pm15 = pow(C1 / T, -1.5);
  t.eH = exp(Uk0 + H.el_ionpot[0] / kT);
  return t;
}

// Stage factors of the LTE ionisation fractions of the group lane's
// elements at temperature T: r = exp(U_j - U_{j-1} - chi_{j-1}/kT) per
// stage j >= 1 (getfjk, rh/solvene.c:180-190; U from getKuruczpf = Linear
// on Tpf, rh/solvene.c:202-211).  They depend on T alone, so each layer
// computes them once into the lane's scratch slots (sc, stride ss) and
// every Newton step of every pressure iteration of the layer reads them.
__device__ void hse_stage_factors(const DevModel &M, const DevHse &H, double T, int sub, int lpc,
                                  double *sc, int64_t ss) {
  const double kT = KBOLTZMANN * T;
  const int npf = M.npf;
  int jT = 0, edge = 0;
  double fx = 0.0;
  if (T <= ldc(M.Tpf)[0]) edge = -1;
  else if (T >= ldc(M.Tpf)[npf - 1]) edge = 1;
  else {
    jT = locate_asc(npf, M.Tpf, T);
    fx = (ldc(M.Tpf)[jT + 1] - T) / (ldc(M.Tpf)[jT + 1] - ldc(M.Tpf)[jT]);
  }
  auto pf = [&](const double *yt) -> double {
    if (edge < 0) return yt[0];
    if (edge > 0) return yt[npf - 1];
    return fx * yt[jT] + (1 - fx) * yt[jT + 1];
  };
  int slot = 0;
  for (int e = sub; e < H.nel; e += lpc) {
    const int ns = H.el_nstage[e];
    const double *lp = H.el_logpf + H.el_pf_ptr[e], *ip = H.el_ionpot + H.el_pf_ptr[e];
    double Uk = pf(lp);
    for (int j = 1; j < ns; j++, slot++) {
      const double Ukp1 = pf(lp + j * npf);
      if (slot < H.nslot_stage) sc[slot * ss] = exp(Ukp1 - Uk - ip[j - 1] / kT);
      Uk = Ukp1;
    }
  }
}

// Saha-Boltzmann level sums of one model atom at temperature T, grouped by
// the ionisation difference dZ to the ground level: g[d] = sum over levels
// i >= 1 with dZ_i = d of g_i/g_0 exp(-(E_i - E_0)/kT) (rh/ltepops.c:87-100),
// d < HSE_NDZ, and g[HSE_NDZ] the same sum over the neutral levels before
// the first ionised one (the neutral fraction of rh/chemequil.c:222-236).
// The LTE normalisation is then 1 + sum_d g[d] / (cNe_T)^d for any n_e.
__device__ void hse_level_groups(const DevModel &M, int ia, double T, double *g, int64_t ss) {
  const int l0 = ldc(M.atom_level_ptr)[ia], nl = ldc(M.atom_level_ptr)[ia + 1] - l0;
  double acc[HSE_NDZ + 1];
#pragma unroll
  for (int d = 0; d <= HSE_NDZ; d++) acc[d] = 0.0;
  bool neutral = ldc(M.level_stage)[l0] == 0;
  for (int i = 1; i < nl; i++) {
    const double dE = ldc(M.level_E)[l0 + i] - ldc(M.level_E)[l0];
    const double r = (ldc(M.level_g)[l0 + i] / ldc(M.level_g)[l0]) * exp(-(dE / (KBOLTZMANN * T)));
    const int dZ = ldc(M.level_stage)[l0 + i] - ldc(M.level_stage)[l0];
    if (ldc(M.level_stage)[l0 + i] > 0) neutral = false;
#pragma unroll
    for (int d = 0; d < HSE_NDZ; d++)
      if (dZ == d) acc[d] += r;
    if (neutral) acc[HSE_NDZ] += r;
  }
#pragma unroll
  for (int d = 0; d <= HSE_NDZ; d++) g[d * ss] = acc[d];
}

// 1 + sum_d g[d] / cNe_T^d
__device__ __forceinline__ double hse_level_sum(const double *g, int64_t ss, double cNe_T) {
  const double ci = 1.0 / cNe_T;
  double sum = 1.0, p = 1.0;
#pragma unroll
  for (int d = 0; d < HSE_NDZ; d++) {
    sum += g[d * ss] * p;
    p *= ci;
  }
  return sum;
}

// get_ne(fromscratch = TRUE) with hydrogen in LTE: the starting value from
// hydrogen ionisation with U = 0 (pyrh_background.c:489-502), then Newton
// steps on charge conservation over every element's LTE ionisation
// fractions (getfjk, rh/solvene.c:145-197, with the stage factors of
// hse_stage_factors) and H-.  The group's lanes take elements sub,
// sub + lpc, ...; lane 0 starts the error from n_e/n_Htot, so with lpc = 1
// the sums run in the reference's order.  The quotients by sum1 and n_e
// are products with their reciprocals (a last-place difference).
// CACHED = false: the stage factors are formed in place (one (column,
// depth) point per lane, get_ne_from_nH), and Uk0 is the starting value's
// partition function (0 in hse(), that of neutral hydrogen in Solve_ne,
// rh/solvene.c:78-88).
template <bool CACHED>
__device__ double hse_get_ne(const DevModel &M, const DevHse &H, double T, double nHtot, int sub,
                             int lpc, const double *sc, int64_t ss, const HseT &tc) {
  const double kT = KBOLTZMANN * T;
  const int npf = M.npf;
  int jT = 0, edge = 0;
  double fx = 0.0;
  if (!CACHED) {
    if (T <= ldc(M.Tpf)[0]) edge = -1;
    else if (T >= ldc(M.Tpf)[npf - 1]) edge = 1;
    else {
      jT = locate_asc(npf, M.Tpf, T);
      fx = (ldc(M.Tpf)[jT + 1] - T) / (ldc(M.Tpf)[jT + 1] - ldc(M.Tpf)[jT]);
    }
  }
  auto pf = [&](const double *yt) -> double {
    if (edge < 0) return yt[0];
    if (edge > 0) return yt[npf - 1];
    return fx * yt[jT] + (1 - fx) * yt[jT + 1];
  };
  const double PhiH = 0.5 * tc.p15 * tc.eH;
  double ne_old = (sqrt(1.0 + 4.0 * nHtot * PhiH) - 1.0) / (2.0 * PhiH), ne = ne_old;
  const double PhiHmin = 0.25 * tc.p15 * tc.eHmin;
  const double CTp = tc.pm15;
  for (int niter = 0; niter < HSE_N_MAX_ELECTRON_ITER; niter++) {
    const double CT_ne = 2.0 * CTp / ne_old, inv_ne = 1.0 / ne_old;
    double err = (sub == 0) ? ne_old / nHtot : 0.0, sum = 0.0;
    int slot = 0;
    for (int e = sub; e < H.nel; e += lpc) {
      const int ns = H.el_nstage[e];
      double u[NSTAGE_MAX], du[NSTAGE_MAX];
      u[0] = 1.0;
      du[0] = 0.0;
      double sum1 = 1.0, sum2 = 0.0;
      const double *lp = H.el_logpf + H.el_pf_ptr[e], *ip = H.el_ionpot + H.el_pf_ptr[e];
      double Uk = CACHED ? 0.0 : pf(lp);
#pragma unroll
      for (int j = 1; j < NSTAGE_MAX; j++) {
        u[j] = du[j] = 0.0;
        if (j < ns) {
          double r;
          if (CACHED) {
            r = sc[(slot + j - 1) * ss];
          } else {
            const double Ukp1 = pf(lp + j * npf);
            r = exp(Ukp1 - Uk - ip[j - 1] / kT);
            Uk = Ukp1;
          }
          u[j] = u[j - 1] * CT_ne * r;
          du[j] = (double) (-j) * u[j] * inv_ne;
          sum1 += u[j];
          sum2 += du[j];
        }
      }
      slot += ns - 1;
      const double inv1 = 1.0 / sum1;
      if (e == 0) {
        const double f0 = u[0] * inv1, df0 = (du[0] - f0 * sum2) * inv1;
        err += ne_old * f0 * PhiHmin;
        sum -= (f0 + ne_old * df0) * PhiHmin;
      }
      const double ab = H.el_abund[e];
#pragma unroll
      for (int j = 1; j < NSTAGE_MAX; j++) {
        if (j < ns) {
          const double f = u[j] * inv1, df = (du[j] - f * sum2) * inv1, akj = ab * j;
          err -= akj * f;
          sum += akj * df;
        }
      }
    }
    // butterfly over the group: every lane ends with the same sums
    for (int o = lpc >> 1; o >= 1; o >>= 1) {
      err += __shfl_xor(err, o);
      sum += __shfl_xor(sum, o);
    }
    ne = ne_old - nHtot * err / (1.0 - nHtot * sum);
    const double dne = fabs((ne - ne_old) / ne_old);
    ne_old = ne;
    if (dne <= HSE_MAX_ELECTRON_ERROR) break;
  }
  return ne;
}

// Continuum opacity at wavelength ls in one layer, summed as
// pyrh_Background does (pyrh_background.c:256-400): H- bound-free and
// free-free times the H- fudge factor, OH and CH bound-free, H bound-free
// and free-free, H2+ and H2- free-free, then the scattering (Thomson,
// Rayleigh on H, He, H2) times its fudge factor.  No metal bound-free.
// hpop(i): population of hydrogen level i.
template <typename HP>
__device__ double hse_opacity(const DevModel &M, const DevHse &H, int ls, double T, double ne,
                              double nH0, double np, double nHmin, double nH2, double nOH,
                              double nCH, double nHe0, HP hpop) {
  const double *lc = M.lamconst + (int64_t) ls * RHG_LC_NCONST;
  const double C0 = Q_ELECTRON * Q_ELECTRON / (4.0 * PI * EPSILON_0) / sqrt(M_ELECTRON);
  const double sigma_ff = 4.0 / 3.0 * sqrt(2.0 * PI / (3.0 * KBOLTZMANN)) * (C0 * C0 * C0) /
                          (HPLANCK * CLIGHT);
  double total = 0.0, chi, stim, sca = ne * H.sigma_thomson;
  if (lc[RHG_LC_HMIN_BF_ON] != 0.0) {
    stim = exp(-lc[RHG_LC_HC_KLA] / T);
    total += nHmin * (1.0 - stim) * lc[RHG_LC_HMIN_BF_ALPHA];
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
  total += chi;
  total *= H.hmin_fudge;
  if (lc[RHG_LC_OH_EIDX] >= 0.0) {
    if (T < ldc(M.oh_T)[0] || T > ldc(M.oh_T)[NTOH - 1]) chi = 0.0;
    else {
      int i2 = locate_asc(NTOH, M.oh_T, T);
      double ti = (i2 == NTOH - 1) ? (double) i2
                                   : (double) i2 + (T - ldc(M.oh_T)[i2]) / (ldc(M.oh_T)[i2 + 1] - ldc(M.oh_T)[i2]);
      double kappa = exp(LG10 * bilinear(NTOH, NEOH, M.oh_cross, ti, lc[RHG_LC_OH_EIDX])) *
                     (CM_TO_M * CM_TO_M);
      stim = exp(-lc[RHG_LC_HC_KLA] / T);
      chi = nOH * (1.0 - stim) * kappa;
    }
    total += chi;
  }
  if (lc[RHG_LC_CH_EIDX] >= 0.0) {
    if (T < ldc(M.ch_T)[0] || T > ldc(M.ch_T)[NTCH - 1]) chi = 0.0;
    else {
      int i2 = locate_asc(NTCH, M.ch_T, T);
      double ti = (i2 == NTCH - 1) ? (double) i2
                                   : (double) i2 + (T - ldc(M.ch_T)[i2]) / (ldc(M.ch_T)[i2 + 1] - ldc(M.ch_T)[i2]);
      double kappa = exp(LG10 * bilinear(NTCH, NECH, M.ch_cross, ti, lc[RHG_LC_CH_EIDX])) *
                     (CM_TO_M * CM_TO_M);
      stim = exp(-lc[RHG_LC_HC_KLA] / T);
      chi = nCH * (1.0 - stim) * kappa;
    }
    total += chi;
  }
  {  // H bound-free, rh/hydrogen.c:150-208 (active-layer branch)
    const int cp0 = ldc(M.lam_cont_ptr)[ls], cp1 = ldc(M.lam_cont_ptr)[ls + 1];
    bool any = false;
    chi = 0.0;
    for (int e = cp0; e < cp1; e++) {
      const int ci = ldc(M.lam_cont_idx)[e];
      if (ldc(M.cont_atom)[ci] != 0) continue;
      any = true;
      const double expla = exp(-lc[RHG_LC_BF_HC_KLA] / T);
      chi += ldc(M.lam_cont_alpha)[e] * (1.0 - expla) * hpop(ldc(M.cont_i)[ci]);
    }
    if (any) total += chi;
  }
  {
    double y = (lc[RHG_LC_GFF_CY] * T) / (HPLANCK * CLIGHT), x3 = lc[RHG_LC_GFF_X3];
    double gIII = 1.0 + 0.1728 * x3 * (1.0 + y) - 0.0496 * (x3 * x3) * (1.0 + (1.0 + y) * 0.33333333 * y);
    double g_ff = (gIII > 1.0) ? gIII : 1.0;
    stim = 1.0 - exp(-lc[RHG_LC_HC_KLA] / T);
    total += sigma_ff / sqrt(T) * lc[RHG_LC_HFF_NU3] * ne * np * stim * g_ff;
  }
  if (H.ray_H > 0.0) sca += H.ray_H * nH0;
  if (H.he_atom >= 0 && H.ray_He > 0.0) sca += H.ray_He * nHe0;
  if (lc[RHG_LC_H2P_LIDX] >= 0.0) {
    double ti = clamp_index(NTEMP_H2P, M.h2pff_temp, T);
    double kappa = bilinear(NTEMP_H2P, NFF_H2P, M.h2pff_kappa, ti, lc[RHG_LC_H2P_LIDX]);
    total += (nH0 * 1.0E-29) * (np * 1.0E-20) * kappa;
  }
  if (H.ray_H2 > 0.0) sca += H.ray_H2 * nH2;
  if (lc[RHG_LC_H2M_LIDX] >= 0.0) {
    if (nH2 > 0.0) {
      double ti = clamp_index(NTHETA_H2, M.h2mff_theta, THETA0 / T);
      double pe = ne * KBOLTZMANN * T;
      double kappa = bilinear(NTHETA_H2, NFF_H2, M.h2mff_kappa, ti, lc[RHG_LC_H2M_LIDX]);
      chi = (nH2 * 1.0E-29) * pe * kappa;
    } else chi = 0.0;
    total += chi;
  }
  sca *= H.scatt_fudge;
  return total + sca;
}

// One pass of the layer loop body (pyrh_hse.c:226-238, :295-307): n_e, and
// when the pressure step needs it (tau500 scale) the LTE populations
// (SetLTEQuantities, rh/ltepops.c:33-113 for the atoms that matter here:
// the molecule nuclei, hydrogen, helium), chemical equilibrium and the
// opacity.
template <int NN>
__device__ void hse_layer(const DevModel &M, const DevHse &H, double T, double nHtot, int sub,
                          int lpc, bool opacity, bool first, const double *sc, int64_t ss,
                          const HseT &tc, double *lds_phi, double *lds_nmol, double *lds_saha,
                          int lstride, double &ne_out, double &opac_out) {
  const double ne = hse_get_ne<true>(M, H, T, nHtot, sub, lpc, sc, ss, tc);
  ne_out = ne;
  if (!opacity) return;
  const double cNe_T = 0.5 * ne * tc.p15;
  // level groups of the nuclei (and He) at this layer's T, after the
  // stage factors in the lane's slots
  const double *lg = sc + (int64_t) H.nslot_stage * ss;
  double nstar0n[NN], fn0[NN];
#pragma unroll
  for (int i = 0; i < NN; i++) {
    const int ia = ldc(M.nucl_atom)[i];
    const double *g = lg + (int64_t) i * (HSE_NDZ + 1) * ss;
    const double ntot = ldc(M.atom_abund)[ia] * nHtot;
    nstar0n[i] = ntot / hse_level_sum(g, ss, cNe_T);
    fn0[i] = (nstar0n[i] + g[HSE_NDZ * ss] * nstar0n[i]) / ntot;
  }
  double nnuc[NN], PhiHmin;
  chem_eq<NN>(M, T, ne, nHtot, fn0, nnuc, lds_phi, lds_nmol, lds_saha, lstride, PhiHmin, !first,
              &tc.p15);
  double fracn[NN];
#pragma unroll
  for (int i = 0; i < NN; i++) fracn[i] = nnuc[i] / (ldc(M.atom_abund)[ldc(M.nucl_atom)[i]] * nHtot);
  const int lH = ldc(M.atom_level_ptr)[0], nlH = ldc(M.atom_level_ptr)[1] - lH;
  const double nH0 = nstar0n[0] * fracn[0];
  const double np = (level_ratio(M, lH, nlH - 1, T, cNe_T) * nstar0n[0]) * fracn[0];
  const double nHmin = ne * (nnuc[0] * PhiHmin);
  const double nH2 = M.mol_h2 >= 0 ? lds_nmol[M.mol_h2 * lstride] : 0.0;
  const double nOH = M.mol_oh >= 0 ? lds_nmol[M.mol_oh * lstride] : 0.0;
  const double nCH = M.mol_ch >= 0 ? lds_nmol[M.mol_ch * lstride] : 0.0;
  double nHe0 = 0.0;
  if (H.he_atom >= 0 && H.ray_He > 0.0) {
    const int ia = H.he_atom;
    nHe0 = (ldc(M.atom_abund)[ia] * nHtot) / hse_level_sum(lg + (int64_t) NN * (HSE_NDZ + 1) * ss, ss,
                                                       cNe_T);
#pragma unroll
    for (int i = 0; i < NN; i++)
      if (ldc(M.nucl_atom)[i] == ia) nHe0 = nstar0n[i] * fracn[i];
  }
  const double n0H = nstar0n[0], frH = fracn[0];
  auto hpop = [&](int i) -> double {
    return ((i == 0) ? n0H : level_ratio(M, lH, i, T, cNe_T) * n0H) * frH;
  };
  opac_out = hse_opacity(M, H, H.lam_index, T, ne, nH0, np, nHmin, nH2, nOH, nCH, nHe0, hpop);
}

}  // namespace

template <int NN>
__global__ void __launch_bounds__(HSE_BLOCK, 2)
k_hse(const DevModel *__restrict__ Mp, DevHse H, int lpc, int ncol, int Nmax,
      const int32_t *__restrict__ ndep, int atm_scale, const double *__restrict__ scale,
      const double *__restrict__ temp, const double *__restrict__ pg_top, double *__restrict__ ne_out,
      double *__restrict__ nH_out, double *__restrict__ rho_out, double *__restrict__ pg_out,
      int32_t *__restrict__ niter_out) {
  const DevModel &M = *Mp;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int ic = gid / lpc, sub = gid % lpc;
  if (ic >= ncol) return;
  extern __shared__ double lds_mol[];
  const int lstride = blockDim.x;
  double *lds_phi = lds_mol + threadIdx.x;
  double *lds_nmol = lds_phi + M.nmol * lstride;
  double *lds_saha = lds_nmol + M.nmol * lstride;
  const int N = ndep[ic];
  const int64_t o = (int64_t) ic * Nmax;
  const double *sc = scale + o, *Tc = temp + o;
  const bool store = (sub == 0);
  const bool tau = (atm_scale == 0);
  const double g = M.gravity, LOG10 = log(10.0), wH = AMU * M.wght_per_H;
  const double tA = M.totalAbund;

  // this lane's stage-factor slots
  const int64_t ss = (int64_t) gridDim.x * blockDim.x;
  double *sc_t = H.scr + gid;

  // top layer, pyrh_hse.c:210-273
  double T = Tc[0], pg = pg_top[ic], ne = 0.0, opac = 0.0, rho = 0.0;
  HseT tc;
  auto layer_tables = [&](double Tl) {
    tc = hse_tcache(H, Tl, 0.0);
    hse_stage_factors(M, H, Tl, sub, lpc, sc_t, ss);
    if (!tau) return;
    double *lg = sc_t + (int64_t) H.nslot_stage * ss;
#pragma unroll
    for (int i = 0; i < NN; i++)
      hse_level_groups(M, ldc(M.nucl_atom)[i], Tl, lg + (int64_t) i * (HSE_NDZ + 1) * ss, ss);
    if (H.he_atom >= 0 && H.ray_He > 0.0)
      hse_level_groups(M, H.he_atom, Tl, lg + (int64_t) NN * (HSE_NDZ + 1) * ss, ss);
  };
  layer_tables(T);
  double nH = (pg / KBOLTZMANN / T - ne) / tA;
  int iter = 0;
  while (iter < H.nmax_iter) {
    rho = wH * nH;
    hse_layer<NN>(M, H, T, nH, sub, lpc, tau, iter == 0, sc_t, ss, tc, lds_phi, lds_nmol,
                  lds_saha, lstride, ne, opac);
    const double old = nH;
    nH = (pg / KBOLTZMANN / T - ne) / tA;
    iter++;
    if (fabs((nH - old) / nH) <= H.iter_limit) break;
  }
  if (store) {
    ne_out[o] = ne;
    nH_out[o] = nH;
    rho_out[o] = rho;
    pg_out[o] = pg;
    if (niter_out) niter_out[o] = iter;
  }
  double z_prev = tau ? pow10d(sc[0]) : sc[0] * KM_TO_M;
  // the layers below, pyrh_hse.c:278-351
  for (int k = 1; k < N; k++) {
    const double z = tau ? pow10d(sc[k]) : sc[k] * KM_TO_M;
    const double pg_prev = pg, rho_prev = rho, opac_prev = opac, ne_prev = ne;
    double deltaP;
    if (tau) deltaP = g * rho_prev / opac_prev * (z - z_prev);
    else deltaP = g * rho_prev * (z_prev - z);
    T = Tc[k];
    layer_tables(T);
    pg = pg_prev + deltaP;
    nH = (pg / KBOLTZMANN / T - ne_prev) / tA;
    const double dlogtau = tau ? log10(z) - log10(z_prev) : 0.0;
    iter = 0;
    while (iter < H.nmax_iter) {
      rho = wH * nH;
      hse_layer<NN>(M, H, T, nH, sub, lpc, tau, iter == 0, sc_t, ss, tc, lds_phi, lds_nmol,
                  lds_saha, lstride, ne, opac);
      if (tau) {
        const double beta1 = rho_prev / opac_prev * z_prev;
        const double beta2 = rho / opac * z;
        pg = pg_prev + LOG10 * g * (beta2 + beta1) / 2 * dlogtau;
      } else {
        pg = pg_prev + g * (z_prev - z) * sqrt(rho * rho_prev);
      }
      const double old = nH;
      nH = (pg / KBOLTZMANN / T - ne) / tA;
      iter++;
      if (fabs((nH - old) / nH) <= H.iter_limit) break;
    }
    if (store) {
      ne_out[o + k] = ne;
      nH_out[o + k] = nH;
      rho_out[o + k] = rho;
      pg_out[o + k] = pg;
      if (niter_out) niter_out[o + k] = iter;
    }
    z_prev = z;
  }
  if (store) {
    for (int k = N; k < Nmax; k++) {
      ne_out[o + k] = 0.0;
      nH_out[o + k] = 0.0;
      rho_out[o + k] = 0.0;
      pg_out[o + k] = 0.0;
      if (niter_out) niter_out[o + k] = 0;
    }
  }
}

// get_ne_from_nH (rh/rhf1d/pyrh_hse.c:555-678): n_e at every (column,
// depth) point from T and n_Htot in cm^-3 (Solve_ne from scratch,
// rh/solvene.c:53-135), one point per lane; n_e returned in cm^-3 as the
// entry point converts back (and zero past ndep).
__global__ void __launch_bounds__(HSE_BLOCK)
k_ne_from_nh(const DevModel *__restrict__ Mp, DevHse H, int ncol, int Nmax,
             const int32_t *__restrict__ ndep, const double *__restrict__ temp,
             const double *__restrict__ nH, double *__restrict__ ne_out) {
  const DevModel &M = *Mp;
  const int64_t gid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (int64_t) ncol * Nmax) return;
  const int ic = (int) (gid / Nmax), k = (int) (gid % Nmax);
  if (k >= ndep[ic]) {
    ne_out[gid] = 0.0;
    return;
  }
  const double CUBE_CM = CM_TO_M * CM_TO_M * CM_TO_M;
  const double T = temp[gid], nHtot = nH[gid] / CUBE_CM;
  const double Uk = linear1(M.npf, M.Tpf, H.el_logpf, T);
  ne_out[gid] = hse_get_ne<false>(M, H, T, nHtot, 0, 1, nullptr, 0, hse_tcache(H, T, Uk)) * CUBE_CM;
}

#define RHG_INST_HSE(NN)                                                                         \
  template __global__ void k_hse<NN>(const DevModel *, DevHse, int, int, int, const int32_t *, int, \
                                     const double *, const double *, const double *, double *,  \
                                     double *, double *, double *, int32_t *);
RHG_INST_HSE(1)
RHG_INST_HSE(2)
RHG_INST_HSE(3)
RHG_INST_HSE(4)
RHG_INST_HSE(5)
RHG_INST_HSE(6)
RHG_INST_HSE(7)
RHG_INST_HSE(8)
#undef RHG_INST_HSE

}  // namespace rhg
