/* Driver for the reference RH library compiled from source (oracle/_ref).

   Test infrastructure only: it runs pyrh's C entry point rhf1d()
   (rh/rhf1d/pyrh_compute1dray.c:112) on a batch of columns read from a
   binary file and writes the emergent Stokes vectors, so that golden
   vectors can be produced in a container that holds the reference sources.

   Optionally it records per-stage intermediates by interposing on the
   library's exported functions (the executable is linked with
   -rdynamic, and calls between translation units of the -fPIC library go
   through the PLT, so these wrappers see every call and forward it to the
   real function obtained with dlsym(RTLD_NEXT)).  Nothing in the library
   is modified.

   Usage:  rhref_run <rundir> <in.bin> <out.bin> [capture.bin]
           rhref_run --hse <rundir> <in.bin> <out.bin>

   rhref_run --nefromnh <rundir> <in.bin> <out.bin>: pyrh's get_ne_from_nH
   (rh/rhf1d/pyrh_hse.c:555) per column from rows 0, 1 and 8; out.bin per
   column float64 ne[Ndep] in cm^-3.
   rhref_run --scales <rundir> <in.bin> <out.bin>: pyrh's get_scales
   (rh/rhf1d/pyrh_hse.c:402) per column with lam_ref in the mu field;
   out.bin per column float64 tau[Ndep], height[Ndep], cmass[Ndep].
   hse mode: in.bin as above (only rows 0 and 1 of each column are used,
   row 0 = scale, row 1 = T; the wave list is ignored, mu holds pg_top;
   overrides "abun id value" and "fudge lambda v0 v1 v2" are passed on),
   out.bin per column float64 ne[Ndep], nHtot[Ndep], rho[Ndep], pg[Ndep]
   from pyrh's hse() (rh/rhf1d/pyrh_hse.c:67).

   in.bin  : int32 magic 0x4e494852 ("RHIN"), Ncol, Ndep, Nwave, atm_scale,
             float64 mu, float64 wave[Nwave], float64 atmos[Ncol][9][Ndep]
   out.bin : per column int32 nlw, float64 lam[nlw], I[nlw], Q[nlw],
             U[nlw], V[nlw]  (Q/U/V are zeros when the library returns none)
   capture : records  char tag[16], int32 col, nspect, mu, to_obs, n,
             float64 data[n]
*/
#define _GNU_SOURCE
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "rh.h"
#include "atom.h"
#include "atmos.h"
#include "rhf1d/geometry.h"
#include "spectrum.h"
#include "background.h"
#include "rhf1d/pyrh_compute1dray.h"
#include "inputs.h"

void get_ne_from_nH(char *cwd, int atm_scale, int Ndep, double *scale, double *temp,
                    double *nH, double *ne);
void get_scales(char *cwd, int Ndep, double *scale, double *temp, double *ne, double *vz,
                double *vmic, double *nH, int atm_scale, double lam_ref, double *tau,
                double *height, double *cmass, int Nabun, int *abundance_id,
                double *abundance_value);
void hse(char *cwd, int Ndep, double *scale, double *temp, double *ne,
         double *nHtot, double *rho, double *pg, int atm_scale,
         int fudge_num, double *fudge_lam, double *fudge,
         int Nabun, int *abundance_id, double *abundance_value);

extern Atmosphere atmos;
extern Geometry geometry;
extern Spectrum spectrum;
extern InputData input;

static FILE *cap = NULL;
static int cur_col = 0;
static int cur_nspect = -1;

static void rec(const char *tag, int nspect, int mu, int to_obs,
                const double *x, int n)
{
  char t[16];
  int hdr[5];
  if (!cap || x == NULL) return;
  memset(t, 0, sizeof(t));
  strncpy(t, tag, 15);
  hdr[0] = cur_col; hdr[1] = nspect; hdr[2] = mu; hdr[3] = to_obs; hdr[4] = n;
  fwrite(t, 1, 16, cap);
  fwrite(hdr, sizeof(int), 5, cap);
  fwrite(x, sizeof(double), n, cap);
}

#define REAL(name) static __typeof__(name) *real_##name = NULL; \
  if (!real_##name) real_##name = (__typeof__(name) *) dlsym(RTLD_NEXT, #name)

/* --- background opacity records, one per (lambda, mu, to_obs) -------- */
int writeBackground(int nspect, int mu, bool_t to_obs, double *chi_c,
                    double *eta_c, double *sca_c, double *chip_c)
{
  REAL(writeBackground);
  int N = atmos.Nspace;
  int nrec = atmos.backgrflags[nspect].ispolarized ? 4 : 1;
  rec("bg_chi", nspect, mu, to_obs, chi_c, nrec * N);
  rec("bg_eta", nspect, mu, to_obs, eta_c, nrec * N);
  rec("bg_sca", nspect, mu, to_obs, sca_c, N);
  return real_writeBackground(nspect, mu, to_obs, chi_c, eta_c, sca_c, chip_c);
}

/* --- storage of the background records (rh/sortlambda.c:564-585).  With
   MAGNETO_OPTICAL the library gives each wavelength's magneto-optical record
   3 * Nspace doubles, but writeBackground copies NrecStokes = 4 records
   into it (rh/readj.c:328), writing Nspace doubles past the block.  This
   allocation gives that record 4 * Nspace doubles, so the copy stays inside
   its block; everything else is as the library allocates it, and the
   library's freeOpacityEmissivity releases it.  Results read only the
   first 3 * Nspace doubles (rh/readj.c:252-262). ------------------------ */
void allocateOpacityEmissivity(void)
{
  int nspect, N = atmos.Nspace;

  spectrum.chi_c_lam = (double **) malloc(spectrum.Nspect * sizeof(double *));
  for (nspect = 0; nspect < spectrum.Nspect; nspect++)
    spectrum.chi_c_lam[nspect] = (double *) malloc(4 * N * sizeof(double));
  if (input.magneto_optical) {
    spectrum.chip_c_lam = (double **) malloc(spectrum.Nspect * sizeof(double *));
    for (nspect = 0; nspect < spectrum.Nspect; nspect++)
      spectrum.chip_c_lam[nspect] = (double *) malloc(4 * N * sizeof(double));
  }
  spectrum.eta_c_lam = (double **) malloc(spectrum.Nspect * sizeof(double *));
  for (nspect = 0; nspect < spectrum.Nspect; nspect++)
    spectrum.eta_c_lam[nspect] = (double *) malloc(4 * N * sizeof(double));
  spectrum.sca_c_lam = (double **) malloc(spectrum.Nspect * sizeof(double *));
  for (nspect = 0; nspect < spectrum.Nspect; nspect++)
    spectrum.sca_c_lam[nspect] = (double *) malloc(4 * N * sizeof(double));
  if (getenv("RHREF_TRACE_ALLOC")) fprintf(stderr, "harness: background records allocated\n");
}

/* --- continuum processes (outputs per lambda) ----------------------- */
#define WRAP_CE(name) \
bool_t name(double lambda, double *chi, double *eta) { \
  REAL(name); bool_t r = real_##name(lambda, chi, eta); \
  if (r && chi) { rec(#name "_c", cur_nspect, 0, 0, chi, atmos.Nspace); \
                  rec(#name "_e", cur_nspect, 0, 0, eta, atmos.Nspace); } \
  return r; }
WRAP_CE(Hminus_bf)
WRAP_CE(Hydrogen_bf)
WRAP_CE(OH_bf_opac)
WRAP_CE(CH_bf_opac)

#define WRAP_C(name) \
bool_t name(double lambda, double *chi) { \
  REAL(name); bool_t r = real_##name(lambda, chi); \
  if (r && chi) rec(#name, cur_nspect, 0, 0, chi, atmos.Nspace); \
  return r; }
WRAP_C(Hminus_ff)
WRAP_C(H2minus_ff)
WRAP_C(H2plus_ff)
WRAP_C(Rayleigh_H2)

void Hydrogen_ff(double lambda, double *chi)
{
  REAL(Hydrogen_ff);
  real_Hydrogen_ff(lambda, chi);
  rec("Hydrogen_ff", cur_nspect, 0, 0, chi, atmos.Nspace);
}

bool_t Rayleigh(double lambda, Atom *atom, double *chi)
{
  REAL(Rayleigh);
  bool_t r = real_Rayleigh(lambda, atom, chi);
  if (r) rec(atom == atmos.H ? "Rayleigh_H" : "Rayleigh_He",
             cur_nspect, 0, 0, chi, atmos.Nspace);
  return r;
}

void Thomson(double *chi)
{
  REAL(Thomson);
  real_Thomson(chi);
  rec("Thomson", -1, 0, 0, chi, atmos.Nspace);
}

bool_t Metal_bf(double lambda, int Nmetal, Atom *metals, double *chi,
                double *eta)
{
  REAL(Metal_bf);
  bool_t r = real_Metal_bf(lambda, Nmetal, metals, chi, eta);
  rec("Metal_bf_c", cur_nspect, 0, 0, chi, atmos.Nspace);
  rec("Metal_bf_e", cur_nspect, 0, 0, eta, atmos.Nspace);
  return r;
}

/* Planck is called once per wavelength at the top of the Background loop
   (rh/background.c:334); it is used here to learn the current index.    */
void Planck(int Nspace, double *T, double lambda, double *Bnu,
            int active_layer)
{
  REAL(Planck);
  int n;
  for (n = 0; n < spectrum.Nspect; n++)
    if (spectrum.lambda && spectrum.lambda[n] == lambda) { cur_nspect = n; break; }
  real_Planck(Nspace, T, lambda, Bnu, active_layer);
}

flags rlk_opacity(double lambda, int nspect, int mu, bool_t to_obs,
                  double *chi, double *eta, double *scatt, double *chip)
{
  REAL(rlk_opacity);
  flags f = real_rlk_opacity(lambda, nspect, mu, to_obs, chi, eta, scatt, chip);
  if (f.hasline) {
    rec("rlk_chi", nspect, mu, to_obs, chi, (atmos.Stokes ? 4 : 1) * atmos.Nspace);
    rec("rlk_eta", nspect, mu, to_obs, eta, (atmos.Stokes ? 4 : 1) * atmos.Nspace);
  }
  return f;
}

flags passive_bb(double lambda, int nspect, int mu, bool_t to_obs,
                 double *chi, double *eta, double *chip)
{
  REAL(passive_bb);
  flags f = real_passive_bb(lambda, nspect, mu, to_obs, chi, eta, chip);
  if (f.hasline) rec("passive_bb_chi", nspect, mu, to_obs, chi, atmos.Nspace);
  return f;
}

/* Voigt profile evaluations made by the Kurucz-line opacity: records
   (a, v, H, F) when the capture of profiles is switched on. */
static int cap_voigt = 0;
double Voigt(double a, double v, double *F, enum VoigtAlgorithm algorithm)
{
  REAL(Voigt);
  double H = real_Voigt(a, v, F, algorithm), r[4];
  if (cap_voigt) {
    r[0] = a; r[1] = v; r[2] = H; r[3] = F ? *F : 0.0;
    rec("voigt", cur_nspect, algorithm, 0, r, 4);
  }
  return H;
}

void LTEpops_elem(Element *element)
{
  REAL(LTEpops_elem);
  int i;
  real_LTEpops_elem(element);
  for (i = 0; i < element->Nstage; i++)
    rec("elem_n", (int) (element - atmos.elements), i, 0, element->n[i], atmos.Nspace);
}

/* --- equilibria, geometry and formal-solution inputs ----------------- */
void ChemicalEquilibrium(int NmaxIter, double iterLimit)
{
  REAL(ChemicalEquilibrium);
  int n;
  real_ChemicalEquilibrium(NmaxIter, iterLimit);
  rec("nHmin", -1, 0, 0, atmos.nHmin, atmos.Nspace);
  rec("ne", -1, 0, 0, atmos.ne, atmos.Nspace);
  for (n = 0; n < atmos.Nmolecule; n++)
    rec("mol_n", n, 0, 0, atmos.molecules[n].n, atmos.Nspace);
  for (n = 0; n < atmos.H->Nlevel; n++)
    rec("H_n", n, 0, 0, atmos.H->n[n], atmos.Nspace);
  for (n = 0; n < atmos.Natom; n++)
    rec("atom_ntotal", n, 0, 0, atmos.atoms[n].ntotal, atmos.Nspace);
}

void convertScales(Atmosphere *atm, Geometry *geo)
{
  REAL(convertScales);
  real_convertScales(atm, geo);
  rec("height", -1, 0, 0, geo->height, geo->Ndep);
  rec("tau_ref", -1, 0, 0, geo->tau_ref, geo->Ndep);
  rec("cmass", -1, 0, 0, geo->cmass, geo->Ndep);
}

void Piece_Stokes_Bezier3_1D(int nspect, int mu, bool_t to_obs, double *chi,
                             double **S, double **I, double *Psi)
{
  REAL(Piece_Stokes_Bezier3_1D);
  int N = atmos.Nspace, n;
  real_Piece_Stokes_Bezier3_1D(nspect, mu, to_obs, chi, S, I, Psi);
  if (!to_obs) return;
  rec("fs_chi", nspect, mu, to_obs, chi, N);
  for (n = 0; n < 4; n++) rec("fs_S", nspect, mu, n, S[n], N);
  for (n = 0; n < 4; n++) rec("fs_I", nspect, mu, n, I[n], N);
}

double Feautrier(int nspect, int mu, double *chi, double *S, enum FeautrierOrder F_order,
                 double *P, double *Psi)
{
  REAL(Feautrier);
  int N = atmos.Nspace;
  double r = real_Feautrier(nspect, mu, chi, S, F_order, P, Psi);
  rec("fe_chi", nspect, mu, 0, chi, N);
  rec("fe_S", nspect, mu, 0, S, N);
  rec("fe_P", nspect, mu, 0, P, N);
  if (Psi) rec("fe_Psi", nspect, mu, 0, Psi, N);
  return r;
}

void Piecewise_Bezier3_1D(int nspect, int mu, bool_t to_obs, double *chi,
                          double *S, double *I, double *Psi, double **dI)
{
  REAL(Piecewise_Bezier3_1D);
  int N = atmos.Nspace;
  real_Piecewise_Bezier3_1D(nspect, mu, to_obs, chi, S, I, Psi, dI);
  if (!to_obs) return;
  rec("fs1_chi", nspect, mu, to_obs, chi, N);
  rec("fs1_S", nspect, mu, to_obs, S, N);
  rec("fs1_I", nspect, mu, to_obs, I, N);
}

/* --- NLTE: set-up and iteration intermediates ------------------------ */
void CollisionRate(struct Atom *atom, FILE *fp_atom)
{
  REAL(CollisionRate);
  int ij;
  real_CollisionRate(atom, fp_atom);
  if (!atom->active) return;
  for (ij = 0; ij < atom->Nlevel * atom->Nlevel; ij++)
    rec("coll_C", ij, 0, 0, atom->C[ij], atmos.Nspace);
  for (ij = 0; ij < atom->Nlevel; ij++)
    rec("nlte_nstar", ij, 0, 0, atom->nstar[ij], atmos.Nspace);
}

void Damping(AtomicLine *line, double *adamp)
{
  REAL(Damping);
  real_Damping(line, adamp);
  if (line->atom->active) rec("adamp", (int) (line - line->atom->line), 0, 0, adamp, atmos.Nspace);
}

void getProfiles(void)
{
  REAL(getProfiles);
  int n, kr;
  real_getProfiles();
  for (n = 0; n < atmos.Nactiveatom; n++) {
    Atom *a = atmos.activeatoms[n];
    for (kr = 0; kr < a->Nline; kr++) {
      rec("wphi", kr, 0, 0, a->line[kr].wphi, atmos.Nspace);
      rec("phi0", kr, 0, 0, a->line[kr].phi[0], atmos.Nspace);
      rec("line_lambda", kr, 0, 0, a->line[kr].lambda, a->line[kr].Nlambda);
    }
  }
}

void initScatter(void)
{
  REAL(initScatter);
  int n, i;
  for (n = 0; n < atmos.Nactiveatom; n++)
    for (i = 0; i < atmos.activeatoms[n]->Nlevel; i++)
      rec("n_init", i, 0, 0, atmos.activeatoms[n]->n[i], atmos.Nspace);
  real_initScatter();
  if (spectrum.J)
    for (n = 0; n < spectrum.Nspect; n++) rec("J_scat", n, 0, 0, spectrum.J[n], atmos.Nspace);
}

double updatePopulations(int niter)
{
  REAL(updatePopulations);
  int n, ij;
  for (n = 0; n < atmos.Nactiveatom && niter == 1; n++) {
    Atom *a = atmos.activeatoms[n];
    for (ij = 0; ij < a->Nlevel * a->Nlevel; ij++)
      rec("gamma1", ij, 0, 0, a->Gamma[ij], atmos.Nspace);
  }
  double d = real_updatePopulations(niter), r[2];
  r[0] = niter;
  r[1] = d;
  rec("iter", -1, 0, 0, r, 2);
  return d;
}

/* called once the iterations are over (rh/rhf1d/pyrh_compute1dray.c:352) */
void _solveray(double muz, mySpectrum *spec)
{
  REAL(_solveray);
  int n, i;
  for (n = 0; n < atmos.Nactiveatom; n++) {
    Atom *a = atmos.activeatoms[n];
    for (i = 0; i < a->Nlevel; i++) {
      rec("pops_n", n, i, 0, a->n[i], atmos.Nspace);
      rec("pops_nstar", n, i, 0, a->nstar[i], atmos.Nspace);
    }
  }
  rec("lambda", -1, 0, 0, spectrum.lambda, spectrum.Nspect);
  real__solveray(muz, spec);
}

/* ------------------------------------------------------------------- */
/* Optional per-run overrides of pyrh.compute1d's keyword arguments, from
   the text file named by RHREF_OVERRIDES, one per line:
       loggf <line index> <log gf>
       lam   <line index> <delta lambda [nm]>
       abun  <atomic number> <abundance>
       rf    <output path>      (also request get_atomic_rfs; the response
                                 functions [nlw][Nloggf+Nlam] go there)   */
typedef struct {
  int nloggf, nlam, nabun;
  int loggf_ids[64], lam_ids[64], abun_ids[64];
  double loggf_val[64], lam_val[64], abun_val[64];
  int nfudge;                       /* hse(): opacity fudge table          */
  double fudge_lam[64], fudge_row[64][3], fudge_val[3 * 64];
  FILE *rf_out;
} Overrides;

static int read_overrides(const char *path, Overrides *ov) {
  char key[16], line[512], rfpath[400];
  memset(ov, 0, sizeof(*ov));
  if (!path) return 0;
  FILE *f = fopen(path, "r");
  if (!f) { perror("overrides"); return -1; }
  while (fgets(line, sizeof(line), f)) {
    int id;
    double v;
    if (sscanf(line, "%15s", key) != 1) continue;
    if (strcmp(key, "rf") == 0) {
      if (sscanf(line, "%*s %399s", rfpath) == 1) ov->rf_out = fopen(rfpath, "wb");
      if (!ov->rf_out) { perror("rf output"); fclose(f); return -1; }
      continue;
    }
    if (strcmp(key, "fudge") == 0 && ov->nfudge < 64) {
      double *r = ov->fudge_row[ov->nfudge];
      if (sscanf(line, "%*s %lf %lf %lf %lf", &ov->fudge_lam[ov->nfudge], r, r + 1, r + 2) == 4)
        ov->nfudge++;
      continue;
    }
    if (sscanf(line, "%*s %d %lf", &id, &v) != 2) continue;
    if (strcmp(key, "loggf") == 0 && ov->nloggf < 64) {
      ov->loggf_ids[ov->nloggf] = id; ov->loggf_val[ov->nloggf++] = v;
    } else if (strcmp(key, "lam") == 0 && ov->nlam < 64) {
      ov->lam_ids[ov->nlam] = id; ov->lam_val[ov->nlam++] = v;
    } else if (strcmp(key, "abun") == 0 && ov->nabun < 64) {
      ov->abun_ids[ov->nabun] = id; ov->abun_val[ov->nabun++] = v;
    }
  }
  fclose(f);
  for (int n = 0; n < 3; n++)
    for (int k = 0; k < ov->nfudge; k++) ov->fudge_val[n * ov->nfudge + k] = ov->fudge_row[k][n];
  return 0;
}

int main(int argc, char **argv)
{
  int hdr[5], Ncol, Ndep, Nwave, atm_scale, c, n;
  double mu, *wave, *atm, *w;
  FILE *fi, *fo;

  int hse_mode = 0, ne_mode = 0, scales_mode = 0;
  Overrides ov;
  if (argc > 1 && strcmp(argv[1], "--hse") == 0) { hse_mode = 1; argv++; argc--; }
  else if (argc > 1 && strcmp(argv[1], "--nefromnh") == 0) { ne_mode = 1; argv++; argc--; }
  else if (argc > 1 && strcmp(argv[1], "--scales") == 0) { scales_mode = 1; argv++; argc--; }
  if (argc < 4) {
    fprintf(stderr, "usage: %s rundir in.bin out.bin [capture.bin]\n", argv[0]);
    return 2;
  }
  if (chdir(argv[1]) != 0) { perror("chdir"); return 2; }
  fi = fopen(argv[2], "rb");
  if (!fi) { perror("in"); return 2; }
  if (fread(hdr, sizeof(int), 5, fi) != 5 || hdr[0] != 0x4e494852) {
    fprintf(stderr, "bad input header\n");
    return 2;
  }
  Ncol = hdr[1]; Ndep = hdr[2]; Nwave = hdr[3]; atm_scale = hdr[4];
  if (fread(&mu, sizeof(double), 1, fi) != 1) return 2;
  wave = malloc(Nwave * sizeof(double));
  w = malloc(Nwave * sizeof(double));
  atm = malloc((size_t) 9 * Ndep * sizeof(double));
  if (fread(wave, sizeof(double), Nwave, fi) != (size_t) Nwave) return 2;
  fo = fopen(argv[3], "wb");
  if (!fo) { perror("out"); return 2; }
  if (argc > 4) cap = fopen(argv[4], "wb");
  if (getenv("RHREF_CAPTURE_VOIGT")) cap_voigt = 1;
  if (read_overrides(getenv("RHREF_OVERRIDES"), &ov) != 0) return 2;

  for (c = 0; c < Ncol; c++) {
    mySpectrum spec;
    double *zero;
    if (fread(atm, sizeof(double), 9 * Ndep, fi) != (size_t) 9 * Ndep) return 2;
    if (ne_mode) {
      double *o = calloc((size_t) Ndep, sizeof(double));
      get_ne_from_nH(".", atm_scale, Ndep, atm, atm + Ndep, atm + 8 * Ndep, o);
      fwrite(o, sizeof(double), (size_t) Ndep, fo);
      free(o);
      fflush(fo);
      continue;
    }
    if (scales_mode) {
      double *o = calloc(3 * (size_t) Ndep, sizeof(double));
      get_scales(".", Ndep, atm, atm + Ndep, atm + 2 * Ndep, atm + 3 * Ndep, atm + 4 * Ndep,
                 atm + 8 * Ndep, atm_scale, mu, o, o + Ndep, o + 2 * Ndep, ov.nabun,
                 ov.nabun ? ov.abun_ids : NULL, ov.nabun ? ov.abun_val : NULL);
      fwrite(o, sizeof(double), 3 * (size_t) Ndep, fo);
      free(o);
      fflush(fo);
      continue;
    }
    if (hse_mode) {
      double *o = calloc(4 * (size_t) Ndep, sizeof(double));
      o[3 * Ndep] = mu;   /* pg[0] = pg_top on input */
      hse(".", Ndep, atm, atm + Ndep, o, o + Ndep, o + 2 * Ndep, o + 3 * Ndep,
          atm_scale, ov.nfudge, ov.nfudge ? ov.fudge_lam : NULL,
          ov.nfudge ? ov.fudge_val : NULL, ov.nabun, ov.nabun ? ov.abun_ids : NULL,
          ov.nabun ? ov.abun_val : NULL);
      fwrite(o, sizeof(double), 4 * (size_t) Ndep, fo);
      free(o);
      fflush(fo);
      continue;
    }
    memcpy(w, wave, Nwave * sizeof(double));
    cur_col = c;
    cur_nspect = -1;
    spec = rhf1d(".", mu, Ndep, atm, atm + Ndep, atm + 2 * Ndep,
                 atm + 3 * Ndep, atm + 4 * Ndep, atm + 5 * Ndep,
                 atm + 6 * Ndep, atm + 7 * Ndep, atm + 8 * Ndep, atm_scale,
                 Nwave, w, ov.nfudge, ov.nfudge ? ov.fudge_lam : NULL,
                 ov.nfudge ? ov.fudge_val : NULL, ov.nloggf, ov.loggf_ids, ov.loggf_val,
                 ov.nlam, ov.lam_ids, ov.lam_val, ov.nabun, ov.abun_ids, ov.abun_val,
                 ov.rf_out != NULL, 0, 0, ".");
    if (ov.rf_out && spec.rfs) {
      int np = ov.nloggf + ov.nlam;
      for (int i = 0; i < spec.nlw; i++) fwrite(spec.rfs[i], sizeof(double), np, ov.rf_out);
      fflush(ov.rf_out);
    }
    zero = calloc(spec.nlw, sizeof(double));
    fwrite(&spec.nlw, sizeof(int), 1, fo);
    fwrite(spec.lam, sizeof(double), spec.nlw, fo);
    fwrite(spec.sI, sizeof(double), spec.nlw, fo);
    fwrite(spec.stokes ? spec.sQ : zero, sizeof(double), spec.nlw, fo);
    fwrite(spec.stokes ? spec.sU : zero, sizeof(double), spec.nlw, fo);
    fwrite(spec.stokes ? spec.sV : zero, sizeof(double), spec.nlw, fo);
    free(zero);
    free(spec.lam); free(spec.sI); free(spec.sQ); free(spec.sU); free(spec.sV);
    fflush(fo);
  }
  (void) n;
  fclose(fo);
  fclose(fi);
  if (cap) fclose(cap);
  if (ov.rf_out) fclose(ov.rf_out);
  return 0;
}
