// NLTE solution of one active atom per column: device copy of the atom,
// the batch's device arrays and the kernels of the accelerated lambda
// iteration (rh/iterate.c:48-143 as rhf1d() runs it).
#pragma once

#include "rhg_kernels.hpp"

namespace rhg {

constexpr double RBOHR = 5.29177349E-11;       // rh/constant.h
constexpr double E_RYDBERG = 2.1798741E-18;

// transitions active at one wavelength, and levels of the active atom
constexpr int NLTE_MAXSET = 8;
constexpr int NLTE_MAXLEV = 16;
constexpr int NLTE_MAXNG = 4;                  // Ng order
// per-thread scratch rows of k_nlte_feautrier and k_nlte_emergent: total
// opacity, source function and the elimination coefficients F and z (ray r
// of a Feautrier chunk in rows NS_F + 2 r and NS_Z + 2 r)
enum { NS_CHI = 0, NS_S, NS_F, NS_Z };
// rays of k_nlte_feautrier solved together (more rays in chunks of this
// many), and the scratch rows per thread of the two kernels
constexpr int NLTE_FE_RAYS = 5;
constexpr int NLTE_NSLOT = 2 + 2 * NLTE_FE_RAYS;
constexpr int NLTE_NSLOT_EM = 4;
// rays swept together by a k_nlte_bezier thread (more rays are swept in
// chunks of this many)
#ifndef RHG_NLTE_RB
#define RHG_NLTE_RB 3
#endif
constexpr int NLTE_RB = RHG_NLTE_RB;

// Device mirror of rhg_nlte_t, plus two derived tables: the transition of
// each line (ln_tr) and each line's first row in the profile array
// (ln_phi).  Entry e of the quadrature of transition t at grid point
// nblue + la is tr_ptr[t] + la; it also numbers the rate accumulators.
struct DevNlte {
  int nl, atom_index, he_atom, ntrans, nline, ncoll, nrays, nentry, nphi;
  double weight, sigma_thomson;
  const double *E, *g;
  const int32_t *stage;
  const int32_t *tr_type, *tr_i, *tr_j, *tr_nblue, *tr_nlam, *tr_ptr, *tr_line;
  const double *tr_wlam, *tr_alpha;
  const double *ln_lambda0, *ln_Aji, *ln_Bji, *ln_Bij, *ln_Grad, *ln_vdw_cross, *ln_vdw_alpha,
      *ln_vdw_cross_he, *ln_stark_c, *ln_stark_C, *ln_stark_Cm;
  const int32_t *ln_vdw, *ln_stark, *ln_tr, *ln_phi;
  const int32_t *as_ptr, *as_tr, *as_bezier;
  const int32_t *coll_type, *coll_i, *coll_j, *coll_ptr;
  const double *coll_T, *coll_c, *coll_M;
  const double *muz, *wmu, *ray_H, *ray_He, *ray_H2;
};

// Device arrays of one batch of ncol columns.  Every per-depth array is
// column-fastest: x[(row * Nmax + k) * ncol + column], row being a level,
// a line, a level pair (i * nl + j), an output wavelength, ...
struct NlteBufs {
  int ncol, Nmax, nfield, nout;
  const int32_t *ndep;
  double *state;                  // LTE column state, fields RHG_CF_*
  double *nstar, *pops;           // [nl]: LTE and current populations
  double *ntot, *vbroad, *nHe;    // [1]: atom total, Doppler width, He I ground
  double *adamp, *wphi;           // [nline]: damping, 1 / profile normalisation
  double *Cmat, *Gam;             // [nl * nl]: collision and rate matrices
  double *phi;                    // [(nphi * nrays + mu) * 2 + to_obs]: profiles
  double *cchi, *ceta, *csca;     // [nout]: background opacity (with scattering),
                                  //         emissivity and scattering opacity
  double *J;                      // [2 * nout]: mean intensity, previous and new
  int32_t *jpar;                  // per column: which half of J is the previous
  double *acc;                    // [nentry * 2 + s]: rate terms per quadrature point
  double *scr;                    // [NLTE_NSLOT] per k_nlte_feautrier thread
  double *dJ;                     // [nlist] x ncol: max |1 - J_old / J| per wavelength
  double *hist;                   // [(ng_order + 2) * nl]: Ng history
};

__global__ void k_nlte_pre(const DevModel *Mp, const DevNlte *Ap, NlteBufs B);
__global__ void k_nlte_bg(const DevModel *Mp, const DevNlte *Ap, NlteBufs B,
                          const int32_t *out_ls);
__global__ void k_nlte_profile(const DevModel *Mp, const DevNlte *Ap, NlteBufs B);
template <int RB, int MS>
__global__ void k_nlte_bezier(const DevModel *Mp, const DevNlte *Ap, NlteBufs B,
                              const int32_t *act, int nact, const int32_t *out_ls,
                              const int32_t *io_list, int nlist, int il_off, int eval_op);
template <int MS>
__global__ void k_nlte_feautrier(const DevModel *Mp, const DevNlte *Ap, NlteBufs B,
                                 const int32_t *act, int nact, const int32_t *out_ls,
                                 const int32_t *io_list, int nlist, int il_off, int eval_op);
__global__ void k_nlte_flip(NlteBufs B, const int32_t *act, int nact);
__global__ void k_nlte_emergent(const DevModel *Mp, const DevNlte *Ap, NlteBufs B,
                                const int32_t *out_ls, const int32_t *io_list, int nlist,
                                double mu, double *stokes);
__global__ void k_nlte_djmax(NlteBufs B, const int32_t *act, int nact, int nlist,
                             double *djmax);
template <int NLP>
__global__ void k_nlte_stateq(const DevNlte *Ap, NlteBufs B, const int32_t *act, int nact,
                              int isum);
__global__ void k_nlte_ng(const DevNlte *Ap, NlteBufs B, const int32_t *act, int nact,
                          int niter, int norder, int ndelay, int nperiod, double *dpops);
__global__ void k_nlte_out(const DevNlte *Ap, NlteBufs B, const double *src, double *out);

}  // namespace rhg
