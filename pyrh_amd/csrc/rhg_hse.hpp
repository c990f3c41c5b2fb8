// Hydrostatic-equilibrium builder (pyrh's hse(), rh/rhf1d/pyrh_hse.c:67-399)
// on the GPU: the device tables and the kernel.
#pragma once

#include "rhg_kernels.hpp"

namespace rhg {

// Device mirror of rhg_hse_t (element tables in device memory).
struct DevHse {
  int nel;
  const int32_t *el_nstage, *el_pf_ptr;
  const double *el_logpf, *el_ionpot, *el_abund;
  int lam_index, he_atom;
  double ray_H, ray_He, ray_H2, sigma_thomson, hmin_fudge, scatt_fudge;
  int nmax_iter;
  double iter_limit;
  // per-lane scratch of the layer's temperature tables, slot s of thread t
  // at scr[s * (launched threads) + t]: nslot_stage ionisation stage
  // factors, then HSE_NDZ + 1 level groups per molecule nucleus and He
  double *scr;
  int nslot_stage;
};

constexpr int HSE_BLOCK = 64;
// ionisation differences to the ground level the level groups distinguish
constexpr int HSE_NDZ = 4;

// One column per group of `lpc` consecutive lanes (lpc a power of two up
// to 16, groups aligned within the wavefront), serial over depth from the
// top.  The lanes of a group split the element sum of each electron-density
// Newton step and combine it by butterfly; everything else they compute
// alike, and lane 0 of the group stores.  Launch ceil(ncol lpc / HSE_BLOCK)
// workgroups of HSE_BLOCK threads with 3 nmol HSE_BLOCK doubles of LDS
// (chemical equilibrium).
template <int NN>
__global__ void k_hse(const DevModel *Mp, DevHse H, int lpc, int ncol, int Nmax,
                      const int32_t *ndep, int atm_scale, const double *scale, const double *temp,
                      const double *pg_top, double *ne_out, double *nH_out, double *rho_out,
                      double *pg_out, int32_t *niter_out);

__global__ void k_ne_from_nh(const DevModel *Mp, DevHse H, int ncol, int Nmax,
                             const int32_t *ndep, const double *temp, const double *nH,
                             double *ne_out);

}  // namespace rhg
