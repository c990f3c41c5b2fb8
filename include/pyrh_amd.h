/* pyrh_amd C ABI: MI355X (gfx950) engine for pyrh's 1-D formal solution.
 *
 * The reference exposes one per-column entry point,
 *     mySpectrum rhf1d(char *cwd, double mu, int Ndep, double *scale, ...)
 * (rh/rhf1d/pyrh_compute1dray.c:112-389, declared in
 *  rh/rhf1d/pyrh_compute1dray.h:26-36 and bound by pyrh.pyx:621-632), which
 * re-parses every input file per call, mutates its arguments and calls
 * exit() on errors.  This ABI splits that call into
 *   - a model (rhg_model_t) that the host builds once per run directory and
 *     wavelength grid from the same keyword/atom/molecule/line files, and
 *   - batched column launches that take caller-owned buffers.
 * All functions return 0 on success or a negative RHG_E* code; the message
 * of the last failure on the calling thread is available from
 * rhg_last_error().  Inputs are never modified.  A context is used by one
 * host thread at a time and owns one HIP stream.
 *
 * Units of the per-column atmosphere follow pyrh.compute1d
 * (pyrh.pyx:537-668): rows 0 scale (log tau500 | log cmass [g cm^-2] |
 * height [km]), 1 T [K], 2 n_e [cm^-3], 3 v_LOS [km/s], 4 v_mic [km/s],
 * 5 B [G], 6 gamma [rad], 7 chi [rad], 8 n_H,tot [cm^-3].
 */
#ifndef PYRH_AMD_H
#define PYRH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RHG_ABI_VERSION 1

enum rhg_status {
  RHG_OK = 0,
  RHG_EINVAL = -1,      /* bad argument or inconsistent model */
  RHG_EDEVICE = -2,     /* HIP runtime error */
  RHG_ENOMEM = -3,      /* device allocation failed */
  RHG_EUNSUPPORTED = -4 /* option outside what the engine implements */
};

/* Solver used for one wavelength, decided from the background flags
 * exactly as rh/rhf1d/formal.c:78-103 does.  BEZIER becomes FEAUTRIER in a
 * static column (the angle-independent branch, formal.c:287-310);
 * BEZIER_ANGLE does not: a polarizable Kurucz line outside FULL_STOKES still
 * marks the wavelength polarized (pyrh sets atmos.Stokes,
 * rh/rhf1d/pyrh_compute1dray.c:258; rh/kurucz.c:656-659), which keeps it
 * angle-dependent (formal.c:98-101). */
enum rhg_solver {
  RHG_SOLVER_FEAUTRIER = 0,   /* no line at this wavelength                */
  RHG_SOLVER_BEZIER = 1,      /* unpolarized cubic Bezier (I only)         */
  RHG_SOLVER_STOKES = 2,      /* cubic DELO-Bezier, full Stokes            */
  RHG_SOLVER_BEZIER_ANGLE = 3 /* cubic Bezier (I only) in every column     */
};

enum rhg_vdw { RHG_VDW_KURUCZ = 0, RHG_VDW_UNSOLD = 1, RHG_VDW_BARKLEM = 2 };

/* Per-wavelength constants computed on the host (column independent). */
enum rhg_lamconst {
  RHG_LC_LAMBDA = 0,     /* nm (vacuum)                                   */
  RHG_LC_HC_KLA,         /* (h c)/(k 1e-9 lambda)      [K]                */
  RHG_LC_TWOHNU3_C2,     /* (2 h c)/(1e-9 lambda)^3                       */
  RHG_LC_BF_HC_KLA,      /* ((h c)/(k 1e-9))/lambda    (bf routines)      */
  RHG_LC_BF_TWOHNU3_C2,  /* ((2 h c)/1e-27)/lambda^3   (bf routines)      */
  RHG_LC_HMIN_BF_ALPHA,  /* H- bf cross-section [m^2] (0 outside range)   */
  RHG_LC_HMIN_BF_ON,     /* 1 if the H- bf term is present                */
  RHG_LC_HMIN_FF_LIDX,   /* fractional lambda index, -1 = long-wave fit   */
  RHG_LC_HFF_NU3,        /* (1e-9 lambda / c)^3                           */
  RHG_LC_GFF_X3,         /* pow(x, 0.33333333) of the ff Gaunt factor     */
  RHG_LC_GFF_CY,         /* (2 lambda 1e-9) k, ff Gaunt factor y = cy T/hc */
  RHG_LC_H2P_LIDX,       /* H2+ ff fractional lambda index (-1 = off)     */
  RHG_LC_H2M_LIDX,       /* H2- ff fractional lambda index (-1 = off)     */
  RHG_LC_OH_EIDX,        /* OH bf fractional energy index (-1 = off)      */
  RHG_LC_CH_EIDX,        /* CH bf fractional energy index (-1 = off)      */
  RHG_LC_HFF_LONG0,      /* 5 coefficients of the long-wave H- ff fit ... */
  RHG_LC_HMIN_FUDGE = RHG_LC_HFF_LONG0 + 5, /* H- opacity fudge factor (1: none)  */
  RHG_LC_SCATT_FUDGE,    /* scattering-opacity fudge factor (NLTE background)    */
  RHG_LC_ITOP,           /* intensity irradiating the top (ATMOS_ITOP; 0: none)  */
  RHG_LC_NCONST
};

/* Per-(column, depth) quantities produced by the equilibria stage and
 * consumed by the wavelength stage (fields of the column state). */
enum rhg_colfield {
  RHG_CF_T = 0, RHG_CF_NE, RHG_CF_VLOS, RHG_CF_VTURB, RHG_CF_B,
  RHG_CF_COSG, RHG_CF_COS2X, RHG_CF_SIN2X, RHG_CF_NHTOT,
  RHG_CF_NHMIN,    /* H- number density                  */
  RHG_CF_NH0,      /* H ground level (after chem. eq.)   */
  RHG_CF_NP,       /* protons                            */
  RHG_CF_NH2, RHG_CF_NOH, RHG_CF_NCH,
  RHG_CF_HEIGHT,   /* [m], after convertScales           */
  RHG_CF_DZ,       /* |z[k-1] - z[k]| [m], k >= 1        */
  RHG_CF_NFIXED    /* continuum and line terms follow    */
};

typedef struct rhg_model {
  int32_t abi_version;

  /* --- switches (keyword.input, after pyrh's overrides) --------------- */
  int32_t stokes_mode;       /* 0 NO_STOKES .. 3 FULL_STOKES          */
  int32_t magneto_optical;   /* only 0 is implemented                 */
  double vmacro_tresh;       /* [m/s]                                 */
  double vmicro_char;        /* [m/s]                                 */
  double wght_per_H;         /* amu per H nucleus                     */
  double totalAbund;
  double gravity;            /* [m s^-2]                              */

  /* --- wavelength grid ------------------------------------------------ */
  int32_t nspect;            /* wavelengths incl. the reference one   */
  int32_t ref_index;         /* index of lambda_ref                   */
  const double *lamconst;    /* [nspect][RHG_LC_NCONST]               */
  const int32_t *solver;     /* [nspect] rhg_solver                   */
  const int32_t *lam_line_ptr;  /* [nspect+1] CSR into lam_line_idx   */
  const int32_t *lam_line_idx;  /* Kurucz line indices per wavelength */
  const int32_t *lam_cont_ptr;  /* [nspect+1] CSR of bf continua      */
  const int32_t *lam_cont_idx;  /* continuum indices per wavelength   */
  const double *lam_cont_alpha; /* cross-section per CSR entry [m^2]  */

  /* --- atoms for LTE populations (levels of all model atoms) ---------- */
  int32_t natom;
  const int32_t *atom_level_ptr; /* [natom+1] into level arrays        */
  const double *level_E;         /* [nlevel] J                         */
  const double *level_g;
  const int32_t *level_stage;
  const double *atom_abund;      /* [natom]                            */

  /* --- bound-free continua (H first, then metals; in reference order) - */
  int32_t ncont;
  const int32_t *cont_atom;      /* [ncont] atom index (0 = H)          */
  const int32_t *cont_i;         /* lower level (within atom)          */
  const int32_t *cont_j;         /* upper level (within atom)          */

  /* --- chemical equilibrium ------------------------------------------ */
  int32_t nnuclei;               /* elements bound in molecules        */
  const int32_t *nucl_atom;      /* [nnuclei] model atom or -1         */
  const int32_t *nucl_elem;      /* [nnuclei] element index            */
  const double *nucl_abund;      /* [nnuclei]                          */
  int32_t nmol;
  const int32_t *mol_fit;        /* [nmol]                             */
  const int32_t *mol_charge;
  const int32_t *mol_nnuclei;    /* total atoms in molecule            */
  const double *mol_ediss;       /* J                                  */
  const double *mol_tmin, *mol_tmax;
  const int32_t *mol_neqc;       /* [nmol]                             */
  const double *mol_eqc;         /* [nmol][8] coefficients (reversed)  */
  const int32_t *mol_nel;        /* [nmol] constituent count           */
  const int32_t *mol_el_nucl;    /* [nmol][4] nucleus index            */
  const int32_t *mol_el_count;   /* [nmol][4]                          */
  int32_t mol_h2, mol_oh, mol_ch;/* molecule indices or -1             */

  /* --- element partition functions (elements used by lines) ---------- */
  int32_t npf;
  const double *Tpf;             /* [npf]                              */
  int32_t nel;                   /* elements carried                   */
  const int32_t *el_nstage;      /* [nel]                              */
  const int32_t *el_pf_ptr;      /* [nel] offset of [nstage][npf] block */
  const double *el_logpf;        /* natural log pf                     */
  const double *el_ionpot;       /* [nel][nstage_max] J, at el_pf_ptr/npf */
  const double *el_abund;        /* [nel]                              */
  const double *el_weight;       /* [nel] amu                          */

  /* --- Kurucz lines (sorted by wavelength) ---------------------------- */
  int32_t nline;
  const int32_t *line_el;        /* [nline] index into element arrays  */
  const int32_t *line_stage;
  const double *line_lambda0;    /* nm                                 */
  const double *line_Ei;         /* J                                  */
  const double *line_hc_la;      /* (h c)/(lambda0 1e-9)               */
  const double *line_Bijhc4pi;   /* hc/4pi Bij x frac x gi             */
  const double *line_twohnu3_c2; /* Aji/Bji                            */
  const double *line_Grad, *line_GStark, *line_GvdW;
  const int32_t *line_vdw;       /* rhg_vdw                            */
  const double *line_cross, *line_alpha;
  const double *line_larmor;     /* (LARMOR) lambda0                   */
  const int32_t *line_polar;     /* polarizable flag                   */
  const int32_t *line_comp_ptr;  /* [nline+1] Zeeman components        */
  const int32_t *comp_q;
  const double *comp_shift, *comp_strength;

  /* --- static continuum tables ---------------------------------------- */
  const double *hminff_theta;    /* [16]     */
  const double *hminff_kappa;    /* [17*16]  */
  const double *h2mff_theta;     /* [8]      */
  const double *h2mff_kappa;     /* [19*8]   */
  const double *h2pff_temp;      /* [10]     */
  const double *h2pff_kappa;     /* [15*10]  */
  const double *oh_T;            /* [15]     */
  const double *oh_cross;        /* [130*15] */
  const double *ch_T;            /* [15]     */
  const double *ch_cross;        /* [105*15] */
} rhg_model_t;

/* Columns of one batch.  atmos is [ncol][9][ndep_max] in pyrh units;
 * column c uses its first ndep[c] depth points. */
typedef struct rhg_columns {
  int32_t ncol;
  int32_t ndep_max;
  const int32_t *ndep;           /* [ncol]                             */
  const double *atmos;           /* [ncol][9][ndep_max]                */
  int32_t atm_scale;             /* 0 tau500, 1 column mass, 2 height  */
  double mu;
} rhg_columns_t;

/* One active (NLTE) atom with its radiative transitions on the merged
 * wavelength grid of the model it is solved with, as pyrh's rhf1d() sets it
 * up for the iteration (rh/readatom.c:62-615, rh/getlambda.c:42-220,
 * rh/sortlambda.c:42-560, rh/collision.c:450-760, rh/broad.c:34-316,
 * rh/gaussleg.c:30-64); built by pyrh_amd.nlte.build_nlte.  The model passed
 * to rhg_open for the NLTE solve is the background model on the same grid,
 * without the active atom's continua.  Transitions are ordered continua
 * first, then lines, as the active sets list them. */
typedef struct {
  int32_t nlevel;           /* levels incl. the continuum level            */
  int32_t atom_index;       /* the atom in the model's atom list           */
  int32_t he_atom;          /* index of He in the model's atoms, -1: none  */
  const double *E, *g;      /* [nlevel] energies [J], statistical weights  */
  const int32_t *stage;     /* [nlevel] ionisation stage                   */
  double weight;            /* atomic weight [amu]                         */
  int32_t ntrans;
  const int32_t *tr_type;   /* 0 line, 1 continuum                         */
  const int32_t *tr_i, *tr_j;          /* lower, upper level               */
  const int32_t *tr_nblue, *tr_nlam;   /* span of the merged grid          */
  const int32_t *tr_ptr;    /* [ntrans+1] offsets into tr_wlam/tr_alpha    */
  const int32_t *tr_line;   /* line index of a line transition, else -1    */
  const double *tr_wlam;    /* quadrature weight per point of the span:
                               lines in Doppler units (rh/getlambda.c:185),
                               continua in nm (rh/getlambda.c:205)         */
  const double *tr_alpha;   /* continuum cross-section per point [m^2]     */
  int32_t nline;
  const double *ln_lambda0, *ln_Aji, *ln_Bji, *ln_Bij, *ln_Grad;
  const int32_t *ln_vdw;    /* 0 none, 1 Unsold, 2 Barklem                 */
  const double *ln_vdw_cross;     /* Unsold: H+He cross-section; Barklem: the
                                     ABO hydrogen cross-section             */
  const double *ln_vdw_alpha;     /* Barklem: velocity exponent alpha       */
  const double *ln_vdw_cross_he;  /* Barklem: Unsold helium cross-section   */
  const int32_t *ln_stark;  /* 0 none, 1 |c| ne, 2 quadratic (Traving)     */
  const double *ln_stark_c, *ln_stark_C, *ln_stark_Cm;
  const int32_t *as_ptr;    /* [nspect+1] active sets over the model grid  */
  const int32_t *as_tr;     /* transition indices                          */
  const int32_t *as_bezier; /* [nspect] 1: scalar Bezier rays, 0: Feautrier */
  int32_t ncoll;
  const int32_t *coll_type; /* OMEGA CE CI CP CH CH0 CH+ = 0..6             */
  const int32_t *coll_i, *coll_j, *coll_ptr;
  const double *coll_T, *coll_c, *coll_M;  /* table, values, spline M      */
  int32_t nrays;
  const double *muz, *wmu;  /* Gauss-Legendre rays on (0, 1)               */
  const double *ray_H, *ray_He, *ray_H2;   /* [nspect] Rayleigh sigma      */
  double sigma_thomson;
  int32_t nmax_iter, nmax_scatter, ng_order, ng_delay, ng_period, isum;
  double iter_limit;
} rhg_nlte_t;

/* Tables of the hydrostatic-equilibrium builder (pyrh's hse(),
 * rh/rhf1d/pyrh_hse.c:67-399) beyond the model: every element that enters
 * the electron-density Newton step of get_ne (rh/rhf1d/pyrh_background.c:
 * 459-549, rh/solvene.c:145-211), hydrogen first, and the scattering
 * cross-sections at the wavelength the opacity is evaluated at.  Built by
 * pyrh_amd.hse.build_hse.  The model passed to rhg_open is the background
 * model on the grid SortLambda makes of {LAMBDA_REF, 500 nm}, without
 * Kurucz lines and with every atom passive. */
typedef struct {
  int32_t nel;              /* elements, hydrogen first                    */
  const int32_t *el_nstage; /* [nel] ionisation stages                     */
  const int32_t *el_pf_ptr; /* [nel] offset of the element's first stage   */
  const double *el_logpf;   /* ln U per stage on the model's Tpf grid      */
  const double *el_ionpot;  /* ionisation potential per stage [J], same
                               offsets (stride npf per element)            */
  const double *el_abund;   /* [nel] abundance relative to hydrogen        */
  int32_t lam_index;        /* model wavelength of the opacity: the last of
                               the grid (pyrh_Background overwrites its
                               total opacity at every wavelength)          */
  int32_t he_atom;          /* index of He in the model's atoms, -1: none  */
  double ray_H, ray_He, ray_H2;  /* Rayleigh cross-sections there [m^2]    */
  double sigma_thomson;
  double hmin_fudge, scatt_fudge;  /* opacity fudge factors there (1: none) */
  int32_t nmax_iter;        /* per-layer pressure iterations, 50           */
  double iter_limit;        /* relative change of nHtot, 1e-2              */
} rhg_hse_t;

typedef struct rhg_context rhg_context_t;

/* Returns the last error message on this thread ("" if none). */
const char *rhg_last_error(void);

/* Library version/ABI check. */
int32_t rhg_abi_version(void);

/* Create a context on HIP device `device`, uploading the model.  The model
 * arrays are copied; the caller may free them afterwards.  Replaces the
 * per-call file parsing at rh/rhf1d/pyrh_compute1dray.c:130-301. */
int rhg_open(const rhg_model_t *model, int device, rhg_context_t **ctx);
int rhg_close(rhg_context_t *ctx);

/* Arithmetic of rhg_synth_lte.  FP64: every stage in double.  FP32: the
 * line opacity (Voigt and Faraday-Voigt profiles), the continuum sums'
 * storage and the DELO-Bezier recursion in float; equilibria, heights,
 * path lengths, the Bezier/w3 weights (they cancel at small dtau) and
 * the Feautrier recursion stay double. */
enum rhg_precision { RHG_PREC_FP64 = 0, RHG_PREC_FP32 = 1 };

/* LTE synthesis of a batch of columns: the work of one rhf1d() call per
 * column with no active atom (pyrh_compute1dray.c:224-348, formal.c:44-346),
 * emergent Stokes vectors at every wavelength of the model except the
 * reference one.  stokes_out is [ncol][4][nspect-1] float64 (I,Q,U,V),
 * host or device memory (device if `device_ptrs` != 0).  `precision` is an
 * rhg_precision.  Depth counts outside [4, ndep_max] return RHG_EINVAL;
 * with device buffers they are checked on the device (no copy to the host
 * before the launches) and the batch of 128 columns or more runs sorted by
 * depth count (rhg_set_tuning) through a permutation, with the results in
 * the caller's column order.  Synchronous. */
int rhg_synth_lte(rhg_context_t *ctx, const rhg_columns_t *cols,
                  int32_t precision, int32_t device_ptrs, double *stokes_out);

/* Wavelength grids of one context, for batches whose columns ask for
 * different wavelengths (every pyrh compute1d call takes its own `wave`,
 * pyrh.pyx:537-541, merged with LAMBDA_REF by SortLambda,
 * rh/sortlambda.c:42-225).  The context's model is built on the union of
 * the grids; grid g is the list of its output positions (the model's output
 * wavelengths, ascending, reference wavelength dropped)
 * grid_out[grid_ptr[g]] .. grid_out[grid_ptr[g + 1] - 1], strictly
 * increasing.  1 <= ngrid <= RHG_MAX_GRIDS; host arrays, copied.  Replaces
 * the previous grids of the context. */
#define RHG_MAX_GRIDS 32
int rhg_set_grids(rhg_context_t *ctx, int32_t ngrid, const int32_t *grid_ptr,
                  const int32_t *grid_out);

/* rhg_synth_lte for a batch whose column i asks for grid col_grid[i] of
 * rhg_set_grids, all in one launch per stage: the columns run sorted by
 * grid and depth count (every wavefront holds columns of one grid), each
 * solved at its grid's wavelengths only.  stokes_out is rhg_synth_lte's
 * [ncol][4][nspect-1] on the union grid; a column's entries at wavelengths
 * outside its grid are zero.  Each solved entry equals what rhg_synth_lte
 * gives on a context built on that grid alone.  col_grid is host or device
 * memory as the other inputs; grid ids outside [0, ngrid) or depth counts
 * outside [4, ndep_max] return RHG_EINVAL.  Needs ngrid (ndep_max + 1) + 1
 * <= 16384.  Synchronous.  rhg_count_work does not cover a grid batch. */
int rhg_synth_lte_grids(rhg_context_t *ctx, const rhg_columns_t *cols,
                        int32_t precision, int32_t device_ptrs,
                        const int32_t *col_grid, double *stokes_out);

/* rhg_synth_lte without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  cols->atmos,
 * cols->ndep and stokes_out are device pointers; the work buffers come from
 * a fixed device array of RHG_NOALLOC_SCRATCH_BYTES per GPU, held from the
 * library's load until the process ends (its address is looked up by
 * rhg_open).  A batch whose work space (column state, continuum, flags)
 * exceeds it is rejected with RHG_EINVAL and a message giving the largest
 * batch of that shape (about 1000 columns of 82 depths and 112 wavelengths
 * in the FP32 form).  Columns run in the caller's order (no sorting by depth
 * count; the results are the same).  Depth counts outside [4, ndep_max]
 * return RHG_EINVAL.  Synchronous.  Calls that overlap on one GPU (two
 * contexts, two host threads) share the array and race; the caller
 * serialises them. */
#define RHG_NOALLOC_SCRATCH_BYTES (128u << 20)
int rhg_synth_lte_noalloc(rhg_context_t *ctx, const rhg_columns_t *cols,
                          int32_t precision, double *stokes_out);

/* Response functions of the emergent Stokes vector: dS(lambda)/dp(k) for
 * the atmosphere parameters below at every depth k, for the wavelengths
 * solved with the polarized DELO-Bezier solver (zero at the others).  Units
 * are those of the atmosphere rows (K, km/s, G, rad, rad, km/s).  Same
 * columns and Stokes output as rhg_synth_lte; rf_out holds
 *   rf[((((p * ndep_max + k) * 4 + s) * (nspect-1) + l) * ncol + c)]
 * in float for RHG_PREC_FP32 and double for RHG_PREC_FP64 (device memory if
 * device_ptrs).  At most RHG_RF_MAX_COLUMNS columns per call, solved in
 * column chunks whose work buffers fit in half of the free device memory
 * (the forward sweep keeps 13 FP64 scalars per depth step and the reverse
 * sweep 36 point and path-length adjoints per depth and Stokes parameter,
 * for every polarized wavelength of the chunk's columns).  The reference
 * computes only d I / d log gf (rh/kurucz.c:224-260); these have no
 * counterpart there. */
enum rhg_rf_param {
  RHG_RF_T = 0,    /* temperature, atmosphere row 1          */
  RHG_RF_VLOS,     /* line-of-sight velocity, row 3          */
  RHG_RF_B,        /* field strength, row 5                  */
  RHG_RF_GAMMA,    /* field inclination, row 6               */
  RHG_RF_CHI,      /* field azimuth, row 7                   */
  RHG_RF_VMIC,     /* microturbulence, row 4                 */
  RHG_RF_NPARAM
};
#define RHG_RF_MAX_COLUMNS 65536  /* per call; run in column chunks that fit the device */
int rhg_rf_lte(rhg_context_t *ctx, const rhg_columns_t *cols, int32_t precision,
               int32_t device_ptrs, double *stokes_out, void *rf_out);

/* The same for a subset of the parameters: params[0..nparam-1] are
 * distinct rhg_rf_param ids, and rf_out holds
 *   rf[((((i * ndep_max + k) * 4 + s) * (nspect-1) + l) * ncol + c)]
 * for the i-th of them.  rhg_rf_lte is this call with all six in enum
 * order.  Only the requested parameters' state tangents and chain terms are
 * computed.  No counterpart in the reference (see rhg_rf_lte). */
int rhg_rf_lte_params(rhg_context_t *ctx, const rhg_columns_t *cols, int32_t precision,
                      int32_t device_ptrs, int32_t nparam, const int32_t *params,
                      double *stokes_out, void *rf_out);

/* d I / d log gf of the emergent intensity, the reference's analytic atomic
 * response functions: compute1d(get_atomic_rfs=True) (pyrh.pyx:604-660;
 * rh/rhf1d/bezier_1D.c:416-536, rh/kurucz.c:224-260, 696-699,
 * rh/rhf1d/formal.c:59-65).  Parameter p perturbs the log gf of model line
 * par_line[p] (an index into the model's line table; < 0 for a parameter
 * with no response, such as the reference's never-filled wavelength-shift
 * entries).  The response is nonzero only at the wavelengths the scalar
 * cubic Bezier solves (moving columns), as in the reference.  stokes_out is
 * as in rhg_synth_lte; rf_out is [ncol][nspect-1][npar] float64, host or
 * device memory (device if `device_ptrs`); par_line is host memory.  FP64.
 * At most RHG_RF_MAX_COLUMNS columns and RHG_RF_MAX_ATOMIC parameters. */
#define RHG_RF_MAX_ATOMIC 256
int rhg_rf_loggf(rhg_context_t *ctx, const rhg_columns_t *cols, int32_t npar,
                 const int32_t *par_line, int32_t device_ptrs, double *stokes_out,
                 double *rf_out);

/* NLTE populations of one active atom in every column of a batch: the work
 * of rhf1d() with the atom ACTIVE up to the converged populations
 * (rh/rhf1d/rhf1d.c:80-140 through rh/iterate.c:48-143: initSolution with
 * LTE populations and J = 0, rh/initial_xdr.c:276-282; N_MAX_SCATTER
 * passes of background scattering, rh/initscatter.c:60-75; then accelerated
 * lambda iterations, the formal solution of every active wavelength with the
 * diagonal operator, rh/rhf1d/formal.c:44-346, rh/fillgamma.c:82-332,
 * statistical equilibrium, rh/statequil.c:36-216, Ng acceleration,
 * rh/accelerate.c:34-146, until the maximum relative change is below
 * ITER_LIMIT or N_MAX_ITER).  Each column stops iterating when it converges
 * (its own rhf1d() would).  The context is opened with the background model
 * on the atom's merged wavelength grid (pyrh_amd.nlte.build_nlte);
 * columns as in rhg_synth_lte.  pops_out is [ncol][nlevel][ndep_max]
 * float64 (zero past ndep), niter_out [ncol] the iterations done and
 * dpops_out [ncol] the last maximum relative change.  stokes_out (NULL to
 * skip) receives what rhf1d() returns afterwards: J updated with the
 * converged populations (up to N_MAX_SCATTER passes,
 * rh/rhf1d/pyrh_compute1dray.c:332-337), then the emergent intensity of
 * one ray at cols->mu (_solveray, rh/rhf1d/pyrh_solveray.c:82-150),
 * [ncol][4][nspect-1] like rhg_synth_lte's, Q, U, V zero.  nstar_out
 * (NULL to skip) receives the atom's LTE populations in the layout of
 * pops_out (pyrh's Populations.nstar, pyrh.pyx:646-656).  Host memory, or
 * device memory for the outputs and the columns if `device_ptrs`.  FP64.
 * Synchronous. */
int rhg_nlte_solve(rhg_context_t *ctx, const rhg_nlte_t *atom, const rhg_columns_t *cols,
                   int32_t device_ptrs, double *pops_out, int32_t *niter_out,
                   double *dpops_out, double *stokes_out, double *nstar_out);

/* Timings of the last rhg_nlte_solve in ms and its work counts:
 * [0] set-up (equilibria, depth scale, background, profiles, collisions),
 * [1] formal solutions at the wavelengths with a line (Bezier rays) in the
 * scattering passes, [2] the same in the iterations, [3] statistical
 * equilibrium, [4] Ng acceleration, [5] total, [6] formal-solution passes,
 * [7] (column, wavelength) pairs solved, [8] and [9] the continuum-only
 * wavelengths (Feautrier) in the scattering passes and the iterations,
 * [10] and [11] the columns summed over the passes without and with the
 * operator's rate terms (scattering and final J passes; iterations). */
#define RHG_NLTE_NTIMING 12
int rhg_nlte_timings(rhg_context_t *ctx, double *out);

/* Gas pressure in hydrostatic equilibrium for a batch of temperature
 * columns: the work of one pyrh hse() call per column (pyrh.pyx:427-487,
 * rh/rhf1d/pyrh_hse.c:67-399).  Layer by layer from the top: p_g from the
 * layer above (on the tau500 scale the trapezoid in log tau of
 * rho tau / chi_500, pyrh_hse.c:312-318; on the height scale the
 * geometric mean of rho, :320-328), n_e from LTE ionisation of every
 * element (get_ne), LTE populations and chemical equilibrium, the
 * continuum opacity at 500 nm (pyrh_Background, without metal bound-free),
 * and n_Htot = (p_g/kT - n_e)/totalAbund, repeated until n_Htot changes by
 * at most iter_limit or nmax_iter times.  scale and temp are
 * [ncol][ndep_max] (log10 tau500 for atm_scale 0, height in km for 2;
 * column mass, 1, is rejected: the reference leaves its pressure step
 * undefined); pg_top [ncol] the gas pressure at the top [Pa].  Outputs are
 * [ncol][ndep_max] in SI units as pyrh returns them (n_e, n_Htot [m^-3],
 * rho [kg m^-3], p_g [Pa]), zero past ndep; niter_out (NULL to skip) the
 * pressure iterations per layer.  Host memory, or device memory for all
 * arrays except `hse` if `device_ptrs`.  FP64.  Synchronous. */
int rhg_hse(rhg_context_t *ctx, const rhg_hse_t *hse, int32_t ncol, int32_t ndep_max,
            const int32_t *ndep, int32_t atm_scale, const double *scale,
            const double *temp, const double *pg_top, int32_t device_ptrs,
            double *ne_out, double *nHtot_out, double *rho_out, double *pg_out,
            int32_t *niter_out);

/* pyrh's get_ne_from_nH for a batch (pyrh.pyx:396-423,
 * rh/rhf1d/pyrh_hse.c:555-678): n_e at every depth point from the
 * temperature [K] and n_Htot [cm^-3] by the LTE ionisation equilibrium of
 * every element (Solve_ne from scratch with hydrogen in LTE,
 * rh/solvene.c:53-135).  temp, nH and ne_out are [ncol][ndep_max], n_e in
 * cm^-3 as the reference returns it, zero past ndep.  The element tables
 * and the context are those of rhg_hse.  Host memory, or device memory for
 * the arrays if `device_ptrs`.  FP64.  Synchronous. */
int rhg_ne_from_nh(rhg_context_t *ctx, const rhg_hse_t *hse, int32_t ncol, int32_t ndep_max,
                   const int32_t *ndep, const double *temp, const double *nH,
                   int32_t device_ptrs, double *ne_out);

/* pyrh's get_scales for a batch (pyrh.pyx:491-533,
 * rh/rhf1d/pyrh_hse.c:402-553): the background opacity at the model's
 * reference wavelength (the lam_ref asked for; rhg_open with the model of
 * that single wavelength) and convertScales (rh/rhf1d/multiatmos.c:100-177)
 * with all three depth scales kept.  Columns as in rhg_synth_lte (rows 1,
 * 2, 3, 4 and 8 are read; mu is not).  tau_out, height_out [m] and
 * cmass_out [kg m^-2] are [ncol][ndep_max], zero past ndep; the input
 * scale comes back converted to SI, where the reference leaves that output
 * unset.  A model with a Kurucz line at the reference wavelength is
 * rejected (RHG_EUNSUPPORTED).  Host memory, or device memory if
 * `device_ptrs`.  FP64.  Synchronous. */
int rhg_scales(rhg_context_t *ctx, const rhg_columns_t *cols, int32_t device_ptrs,
               double *tau_out, double *height_out, double *cmass_out);

/* Duration of the hydrostatic-equilibrium kernel of the last rhg_hse call
 * in ms (HIP events on the context's stream) and the lanes per column it
 * ran with. */
int rhg_hse_timing(rhg_context_t *ctx, double *ms_out, int32_t *lanes_out);

/* Same as rhg_synth_lte but also returns, per column, the depth scale and
 * the background opacity at the reference wavelength (for testing):
 * height_out/chiref_out are [ncol][ndep_max] float64 host buffers (NULL to
 * skip). */
int rhg_synth_lte_diag(rhg_context_t *ctx, const rhg_columns_t *cols,
                       int32_t precision, double *stokes_out,
                       double *height_out, double *chiref_out);

/* Time the kernels of the last rhg_synth_lte call on the context's
 * stream, in milliseconds (HIP events around each stage); out has
 * RHG_NSTAGE entries: column state (equilibria), geometry, continuum,
 * formal solution. */
#define RHG_NSTAGE 4
int rhg_last_timings(rhg_context_t *ctx, double *ms_out);

/* Launch configuration of a context, chosen by the autotuner
 * (pyrh_amd.autotune), which times the candidates on the caller's batch.
 * No counterpart in the reference.  Every setting gives bit-identical
 * results; they differ in speed only.  Values are 0 or 1:
 *   RHG_TUNE_LINE_TABLE   stage each wavelength's line list in LDS for the
 *                         polarized formal solution (default 1; has no
 *                         effect when the largest list exceeds 16 KB)
 *   RHG_TUNE_COLUMN_SORT  run a batch of uneven depth counts sorted by depth
 *                         count (default 1)
 *   RHG_TUNE_WAVE_ORDER   launch the polarized wavelengths in the order of
 *                         their Zeeman-component counts, most first, rather
 *                         than in wavelength order (default 0)
 * rhg_set_tuning returns RHG_EINVAL for an unknown key or value; setting the
 * wavelength order copies a list to the device (synchronous). */
enum rhg_tune { RHG_TUNE_LINE_TABLE = 0, RHG_TUNE_COLUMN_SORT = 1, RHG_TUNE_WAVE_ORDER = 2 };
int rhg_set_tuning(rhg_context_t *ctx, int32_t key, int32_t value);
int rhg_get_tuning(rhg_context_t *ctx, int32_t key, int32_t *value);

/* Kernel milliseconds of the last rhg_rf_lte / rhg_rf_lte_params call, 5
 * entries: synthesis inputs (state, geometry, continuum), state tangents,
 * forward sweep, reverse sweep, chain to the parameters (with the scalar
 * wavelengths' synthesis). */
int rhg_rf_timings(rhg_context_t *ctx, double *ms_out);

/* Algorithmic work census of the last rhg_synth_lte call on this context
 * (its columns, depth counts and column state must still be alive): sums
 * over emergent rays (column, output wavelength) and depth indices
 * 1..N-1 of the counts below, written to counts_out[RHG_NWORK].  The
 * Humlicek branch of every Zeeman component is replayed from the same
 * damping and Doppler offsets the synthesis used.  It is the input of the
 * roofline's FLOP count; it has no counterpart in the reference. */
enum rhg_work {
  RHG_WORK_STOKES = 0,    /* depth points on polarized DELO-Bezier rays */
  RHG_WORK_BEZIER = 1,    /* ... on scalar Bezier rays */
  RHG_WORK_FEAUTRIER = 2, /* ... on Feautrier rays */
  RHG_WORK_LINE = 3,      /* (line in window, depth point) pairs */
  RHG_WORK_VOIGT_I = 4,   /* Zeeman-component Voigt evaluations, region I */
  RHG_WORK_VOIGT_II = 5,
  RHG_WORK_VOIGT_III = 6,
  RHG_WORK_VOIGT_IV = 7,
  RHG_WORK_ARMSTRONG = 8, /* unpolarized-line Voigt evaluations */
  RHG_WORK_GAUSS = 9,     /* Doppler-profile evaluations (no damping) */
  RHG_WORK_COLDEPTH = 10, /* (column, depth) points of the column state */
  RHG_NWORK = 12
};
int rhg_count_work(rhg_context_t *ctx, int64_t *counts_out);

/* sizeof(rhg_model_t) and sizeof(rhg_columns_t) as compiled, for host
 * bindings to check their struct layout. */
int32_t rhg_model_sizeof(void);
int32_t rhg_columns_sizeof(void);

#ifdef __cplusplus
}
#endif

#endif /* PYRH_AMD_H */
