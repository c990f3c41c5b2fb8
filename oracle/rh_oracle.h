/* Test oracle: CPU restatement of the reference RH algorithm for pyrh's
 * LTE synthesis path (rhf1d with no active atom).  TEST INFRASTRUCTURE
 * ONLY: it is the checker for the GPU library and the CPU baseline of the
 * benchmark; nothing in the product links or calls it.
 *
 * Each function names the reference file:line it follows.  It reads the
 * same host tables the GPU library is given (rhg_model_t), so the model
 * builder is exercised by the comparison with the reference itself.
 */
#ifndef RH_ORACLE_H
#define RH_ORACLE_H

#include "../include/pyrh_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Full LTE synthesis of a batch of columns (per column: equilibria,
 * background, convertScales, formal solution of the emergent ray).
 * stokes_out [ncol][4][nspect-1]; height_out/chiref_out [ncol][ndep_max]
 * may be NULL.  inverse_fp32 != 0 inverts the 4x4 DELO matrix in float as
 * the shipped reference does (rh/bezier_aux.c:235-328); 0 inverts in
 * double.  nthreads <= 0 uses all OpenMP threads. */
int rho_synth_lte(const rhg_model_t *m, const rhg_columns_t *cols,
                  int inverse_fp32, int nthreads, double *stokes_out,
                  double *height_out, double *chiref_out);

/* The reference's analytic d I / d log gf (compute1d(get_atomic_rfs=True),
 * rh/rhf1d/bezier_1D.c:416-536) with the FP64 synthesis: parameter p is
 * model line par_line[p] (< 0: no response).  rf_out [ncol][nspect-1][npar]. */
int rho_rf_loggf(const rhg_model_t *m, const rhg_columns_t *cols, int npar,
                 const int *par_line, int nthreads, double *stokes_out, double *rf_out);

/* NLTE populations of the active atom `na` (accelerated lambda iteration
 * as rhf1d() runs it, rh/iterate.c:48-143): the model is the background
 * model on the merged grid.  pops_out [ncol][nlevel][ndep_max]; niter_out
 * and dpops_out [ncol] get the iterations done and the last max relative
 * change; stokes_out (NULL to skip) [ncol][4][nspect-1] the emergent
 * spectrum at cols->mu afterwards (pyrh's _solveray,
 * rh/rhf1d/pyrh_solveray.c:82-107), Q, U, V zero. */
int rho_nlte_solve(const rhg_model_t *m, const rhg_nlte_t *na, const rhg_columns_t *cols,
                   int nthreads, double *pops_out, int *niter_out, double *dpops_out,
                   double *stokes_out);

/* pyrh's hse() for a batch of columns (rh/rhf1d/pyrh_hse.c:67-351): the
 * model is the background model of pyrh_amd.hse.build_hse, h its element
 * and scattering tables.  scale/temp [ncol][ndep_max], pg_top [ncol];
 * outputs [ncol][ndep_max] in SI units, zero past ndep; niter (NULL to
 * skip) the pressure iterations per layer.  atm_scale 0 (tau500) or 2
 * (height). */
int rho_hse(const rhg_model_t *m, const rhg_hse_t *h, int ncol, int ndep_max,
            const int *ndep, int atm_scale, const double *scale, const double *temp,
            const double *pg_top, int nthreads, double *ne, double *nHtot, double *rho,
            double *pg, int *niter);
/* The same, with a work census (NULL to skip): census[0] the Newton steps
 * of get_ne, census[1] the iterations of the chemical equilibrium, summed
 * over the batch. */
int rho_hse_census(const rhg_model_t *m, const rhg_hse_t *h, int ncol, int ndep_max,
                   const int *ndep, int atm_scale, const double *scale, const double *temp,
                   const double *pg_top, int nthreads, double *ne, double *nHtot, double *rho,
                   double *pg, int *niter, long long *census);

/* pyrh's get_ne_from_nH (rh/rhf1d/pyrh_hse.c:555-678): n_e [cm^-3] per
 * depth from temp [K] and nH [cm^-3], all [ncol][ndep_max]; h as for
 * rho_hse. */
int rho_ne_from_nh(const rhg_model_t *m, const rhg_hse_t *h, int ncol, int ndep_max,
                   const int *ndep, const double *temp, const double *nH, int nthreads,
                   double *ne_out);

/* pyrh's get_scales (rh/rhf1d/pyrh_hse.c:402-553) for a batch of columns
 * on a model whose reference wavelength is the lam_ref asked for: tau500,
 * height [m] and column mass [kg m^-2], [ncol][ndep_max]. */
int rho_scales(const rhg_model_t *m, const rhg_columns_t *cols, int nthreads, double *tau_out,
               double *height_out, double *cmass_out);

/* Single-function entry points for closed-form and invariant tests. */
double rho_voigt_humlicek(double a, double v, double *F);
double rho_voigt_armstrong(double a, double v);
void rho_bezier3_coeffs(double dt, double *out5); /* alpha,beta,gamma,theta,eps */
void rho_w3(double dtau, double *w);
double rho_planck(double T, double lambda_nm);

/* Formal solvers on given opacity and source functions along the ray
 * leaving the top (to_obs), height z [m] decreasing with index, with the
 * thermalised lower boundary built from B2 = {B(T[n-2]), B(T[n-1])}
 * (B2 == NULL: zero incoming intensity).  S4 is [4][n] (source functions
 * of I,Q,U,V), chiQUV is [3][n] (chi_Q,U,V, not divided by chi), chip
 * [3][n] the magneto-optical chip_Q,U,V (NULL: terms off);
 * I4 receives I,Q,U,V at every depth, [4][n]; I receives [n]. */
void rho_formal_stokes(int n, const double *z, double mu, const double *chi,
                       const double *S4, const double *chiQUV, const double *chip,
                       const double *B2, int inverse_fp32, double *I4);
void rho_formal_bezier(int n, const double *z, double mu, const double *chi,
                       const double *S, const double *B2, double *I);
/* The same solver along the inward ray (depth 0 to n-1, incident intensity
 * Itop at the top: 0, or ATMOS_ITOP's value), and along the emergent ray with the reference's
 * d I / d log gf recursion: I holds the inward intensities on entry;
 * dchi, deta are [npar][n]; dI0 receives the emergent responses [npar]. */
void rho_formal_bezier_in(int n, const double *z, double mu, const double *chi,
                          const double *S, double Itop, double *I);
void rho_formal_bezier_rf(int n, const double *z, double mu, const double *chi,
                          const double *S, const double *B2, int npar,
                          const double *dchi, const double *deta, double *I, double *dI0);
double rho_formal_feautrier(int n, const double *z, double mu,
                            const double *chi, const double *S,
                            const double *B2);

#ifdef __cplusplus
}
#endif

#endif
