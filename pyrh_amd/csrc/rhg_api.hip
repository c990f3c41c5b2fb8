// C ABI of the engine (include/pyrh_amd.h): context, model upload and the
// batched LTE synthesis launch.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "rhg_hse.hpp"
#include "rhg_nlte.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHECK(call)                                                             \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(RHG_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  int ensure(size_t n) {
    if (n <= bytes) return 0;
    if (p) (void) hipFree(p);
    p = nullptr;
    bytes = 0;
    if (hipMalloc(&p, n) != hipSuccess) return -1;
    bytes = n;
    return 0;
  }
  ~DevBuf() {
    if (p) (void) hipFree(p);
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// The synthesis without allocation (rhg_synth_lte_noalloc).  Its work
// buffers (sanitised depth counts, an error flag, moving flags, column state,
// continuum) live in `scratch`, a file-scope device array of
// RHG_NOALLOC_SCRATCH_BYTES that each GPU holds from the library's load until
// the process ends.  Host code never names the array in a kernel argument:
// rhg_open looks up its device address on each device it opens
// (noalloc_scratch_init) and every call uses that pointer.  Calls that
// overlap on one GPU share the array and race.
static __device__ __align__(256) unsigned char scratch[RHG_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
// LDS for a staged line list, at most: 16 KB keeps 8 one-wavefront
// workgroups per CU (the register limit of the polarized kernels)
constexpr size_t kLineTabMaxBytes = 16384;
std::atomic<void *> g_scratch_ptr[kMaxDevices];
std::mutex g_scratch_mutex;

// with `device` current; contexts opened from several threads look it up once
int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(RHG_EUNSUPPORTED, "device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device].load()) return 0;
  void *p = nullptr;
  HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(RHG_EDEVICE, "no device address for the no-allocation scratch");
  g_scratch_ptr[device].store(p);
  return 0;
}

}  // namespace

struct rhg_context {
  int device = 0;
  hipStream_t stream = nullptr;
  rhg::DevModel dm{};                 // host copy (device pointers inside)
  const rhg::DevModel *d_model = nullptr;  // the same struct in device memory
  std::vector<void *> allocs;   // model arrays
  int nspect = 0, nout = 0, ncont = 0, nline = 0, natom = 0;
  int32_t *d_out_ls = nullptr;
  // output-wavelength positions by solver class: polarized (STOKES) and
  // scalar (BEZIER / FEAUTRIER)
  int n_pol = 0, n_scal = 0;
  // largest line list of a wavelength (lines, Zeeman components of damped
  // polarizable lines) and the LDS bytes of its staged copy; 0 bytes: the
  // polarized kernels read the lists from global memory
  int tab_maxl = 0, tab_maxc = 0;
  size_t tab_bytes = 0;
  // launch configuration (rhg_set_tuning): LDS line table, sorting of
  // batches of uneven depth counts, order of the polarized wavelengths
  int tune_line_table = 1, tune_column_sort = 1, tune_wave_order = 0;
  std::vector<int32_t> h_pol, h_pol_heavy;
  const int32_t *d_io_pol = nullptr, *d_io_scal = nullptr;
  DevBuf atmos, ndep, state, moving, cchi, ceta, stokes, chiref, census, census_sum;
  // response functions: perturbed inputs/states, tangents, tape, adjoints
  DevBuf rf_atm[2], rf_state[2], rf_cont[4], rf_chiref[2], rf_moving, rf_ndep, rf_dstate,
      rf_dcont[2], rf_dchiref, rf_car, rf_aP, rf_aZ, rf_aB, rf_out;
  // log gf response functions: inward intensities, output, parameter lines
  DevBuf at_iin, at_out, at_par;
  // NLTE: the active atom (packed arrays and its DevNlte), the batch arrays
  // of rhg::NlteBufs, active-column list, per-column reductions, output
  DevBuf nl_atom, nl_dev, nl_nstar, nl_pops, nl_ntot, nl_vbroad, nl_nHe, nl_adamp, nl_wphi,
      nl_Cmat, nl_Gam, nl_phi, nl_csca, nl_J, nl_jpar, nl_acc, nl_scr, nl_dJ, nl_hist, nl_act, nl_red,
      nl_out, nl_iolist, nl_atm, nl_ndep, nl_st;
  // batches of uneven depth counts run sorted by depth count: the caller's
  // columns gathered into these, the outputs scattered back
  DevBuf srt_atm, srt_ndep, srt_perm, srt_out, srt_rf;
  // the synthesis's depth counts as its kernels read them (checked, clamped
  // and, for uneven batches, sorted), the slot -> column permutation, the
  // counting sort's histogram and the out-of-range flag
  DevBuf syn_ndep, syn_perm, syn_hist;
  // wavelength grids (rhg_set_grids): per grid its polarized and its scalar
  // output positions (the lists of all grids concatenated, with offsets
  // [ngrid + 1]) and its range of bound-free tiles [2 ngrid]; the longest
  // list of each kind; per grid a bit mask of its output positions
  // (gr_words 32-bit words each); for a grid batch, the block -> grid table
  // (and a grid batch's host grid ids copied in)
  int ngrid = 0, gr_max_pol = 0, gr_max_scal = 0, gr_words = 0;
  DevBuf gr_pol, gr_pol_ptr, gr_scal, gr_scal_ptr, gr_tile, gr_mask, gr_blk, gr_col;
  double nlte_ms[RHG_NLTE_NTIMING]{};
  // hydrostatic equilibrium: element tables, inputs, outputs, timing
  DevBuf hs_nstage, hs_ptr, hs_logpf, hs_ionpot, hs_abund, hs_in, hs_ndep, hs_out, hs_niter,
      hs_scr;
  float hse_ms = 0.0f;
  bool ref_has_line = false;   // a Kurucz line covers the reference wavelength
  int hse_lpc = 0;
  // host copies of model tables the NLTE checks need
  std::vector<int32_t> h_atom_level_ptr, h_nucl_atom, h_level_stage;
  hipEvent_t ev[RHG_NSTAGE + 2]{};
  float last_ms[RHG_NSTAGE + 1]{};
  float rf_ms[5]{};   // the last response-function call's intervals
  // the last batch, for rhg_count_work
  int last_ncol = 0, last_Nmax = 0, last_nfield = 0;
  double last_mu = 1.0;
  const int32_t *last_ndep = nullptr;
};

namespace {

template <typename T>
int upload(rhg_context *c, const T *src, size_t n, const T **dst) {
  if (n == 0) n = 1;
  void *p = nullptr;
  if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return fail(RHG_ENOMEM, "hipMalloc (model)");
  c->allocs.push_back(p);
  if (src) {
    hipError_t e = hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(RHG_EDEVICE, std::string("hipMemcpy (model): ") + hipGetErrorString(e));
  } else {
    (void) hipMemset(p, 0, n * sizeof(T));
  }
  *dst = (const T *) p;
  return 0;
}

}  // namespace

extern "C" {

const char *rhg_last_error(void) { return g_err.c_str(); }

int32_t rhg_abi_version(void) { return RHG_ABI_VERSION; }

int32_t rhg_model_sizeof(void) { return (int32_t) sizeof(rhg_model_t); }
int32_t rhg_columns_sizeof(void) { return (int32_t) sizeof(rhg_columns_t); }

int rhg_close(rhg_context_t *ctx) {
  if (!ctx) return 0;
  (void) hipSetDevice(ctx->device);
  if (ctx->stream) (void) hipStreamSynchronize(ctx->stream);
  for (void *p : ctx->allocs) (void) hipFree(p);
  if (ctx->d_out_ls) (void) hipFree(ctx->d_out_ls);
  for (auto &e : ctx->ev)
    if (e) (void) hipEventDestroy(e);
  if (ctx->stream) (void) hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

int rhg_open(const rhg_model_t *m, int device, rhg_context_t **out) {
  if (!m || !out) return fail(RHG_EINVAL, "null argument");
  *out = nullptr;
  if (m->abi_version != RHG_ABI_VERSION) return fail(RHG_EINVAL, "model ABI version mismatch");
  // a grid of the reference wavelength alone serves the hydrostatic
  // builder; the synthesis entry points need an output wavelength
  if (m->nspect < 1 || m->ref_index < 0 || m->ref_index >= m->nspect)
    return fail(RHG_EINVAL, "bad wavelength grid");
  if (m->natom < 1 || m->natom > RHG_MAX_ATOM) return fail(RHG_EUNSUPPORTED, "too many model atoms");
  if (m->nnuclei < 1 || m->nnuclei > rhg::NNUC_MAX) return fail(RHG_EUNSUPPORTED, "too many molecule nuclei");
  if (m->nmol > rhg::NMOL_MAX) return fail(RHG_EUNSUPPORTED, "too many molecules");
  for (int i = 0; i < m->nel; i++)
    if (m->el_nstage[i] > rhg::NSTAGE_MAX) return fail(RHG_EUNSUPPORTED, "too many ionisation stages");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(RHG_EDEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(RHG_EINVAL, "bad device index");
  HIPCHECK(hipSetDevice(device));

  rhg_context *c = new rhg_context();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(RHG_EDEVICE, "hipStreamCreate");
  }
  for (auto &e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) { rhg_close(c); return fail(RHG_EDEVICE, "hipEventCreate"); }

  rhg::DevModel &d = c->dm;
  const int ns = m->nspect;
  int rc = 0;
#define UP(field, n) \
  if (!rc) rc = upload(c, m->field, (size_t) (n), &d.field)
  d.stokes_mode = m->stokes_mode;
  d.magneto_optical = m->magneto_optical;
  d.vmacro_tresh = m->vmacro_tresh;
  d.vmicro_char = m->vmicro_char;
  d.wght_per_H = m->wght_per_H;
  d.totalAbund = m->totalAbund;
  d.gravity = m->gravity;
  d.nspect = ns;
  d.ref_index = m->ref_index;
  UP(lamconst, (size_t) ns * RHG_LC_NCONST);
  UP(solver, ns);
  UP(lam_line_ptr, ns + 1);
  UP(lam_line_idx, m->lam_line_ptr[ns]);
  UP(lam_cont_ptr, ns + 1);
  UP(lam_cont_idx, m->lam_cont_ptr[ns]);
  UP(lam_cont_alpha, m->lam_cont_ptr[ns]);
  d.natom = m->natom;
  UP(atom_level_ptr, m->natom + 1);
  UP(level_E, m->atom_level_ptr[m->natom]);
  UP(level_g, m->atom_level_ptr[m->natom]);
  UP(level_stage, m->atom_level_ptr[m->natom]);
  UP(atom_abund, m->natom);
  d.ncont = m->ncont;
  UP(cont_atom, m->ncont);
  UP(cont_i, m->ncont);
  UP(cont_j, m->ncont);
  d.nnuclei = m->nnuclei;
  UP(nucl_atom, m->nnuclei);
  UP(nucl_elem, m->nnuclei);
  UP(nucl_abund, m->nnuclei);
  d.nmol = m->nmol;
  UP(mol_fit, m->nmol);
  UP(mol_charge, m->nmol);
  UP(mol_nnuclei, m->nmol);
  UP(mol_ediss, m->nmol);
  UP(mol_tmin, m->nmol);
  UP(mol_tmax, m->nmol);
  UP(mol_neqc, m->nmol);
  UP(mol_eqc, 8 * m->nmol);
  UP(mol_nel, m->nmol);
  UP(mol_el_nucl, 4 * m->nmol);
  UP(mol_el_count, 4 * m->nmol);
  d.mol_h2 = m->mol_h2;
  d.mol_oh = m->mol_oh;
  d.mol_ch = m->mol_ch;
  d.npf = m->npf;
  UP(Tpf, m->npf);
  d.nel = m->nel;
  size_t npfblk = 1;
  if (m->nel > 0) npfblk = (size_t) m->el_pf_ptr[m->nel - 1] + (size_t) m->el_nstage[m->nel - 1] * m->npf;
  UP(el_nstage, m->nel);
  UP(el_pf_ptr, m->nel);
  UP(el_logpf, npfblk);
  UP(el_ionpot, npfblk);
  UP(el_abund, m->nel);
  UP(el_weight, m->nel);
  d.nline = m->nline;
  UP(line_el, m->nline);
  UP(line_stage, m->nline);
  UP(line_lambda0, m->nline);
  UP(line_Ei, m->nline);
  UP(line_hc_la, m->nline);
  UP(line_Bijhc4pi, m->nline);
  UP(line_twohnu3_c2, m->nline);
  UP(line_Grad, m->nline);
  UP(line_GStark, m->nline);
  UP(line_GvdW, m->nline);
  UP(line_vdw, m->nline);
  UP(line_cross, m->nline);
  UP(line_alpha, m->nline);
  UP(line_larmor, m->nline);
  UP(line_polar, m->nline);
  UP(line_comp_ptr, m->nline + 1);
  UP(comp_q, m->line_comp_ptr[m->nline]);
  UP(comp_shift, m->line_comp_ptr[m->nline]);
  UP(comp_strength, m->line_comp_ptr[m->nline]);
  UP(hminff_theta, 16);
  UP(hminff_kappa, 17 * 16);
  UP(h2mff_theta, 8);
  UP(h2mff_kappa, 19 * 8);
  UP(h2pff_temp, 10);
  UP(h2pff_kappa, 15 * 10);
  UP(oh_T, 15);
  UP(oh_cross, 130 * 15);
  UP(ch_T, 15);
  UP(ch_cross, 105 * 15);
#undef UP
  if (rc) { std::string e = g_err; rhg_close(c); return fail(rc, e); }

  // derived tables: Doppler offsets per (wavelength, line) entry and the
  // bound-free cross-sections tiled over output wavelengths
  std::vector<int32_t> ols;
  for (int i = 0; i < ns; i++)
    if (i != m->ref_index) ols.push_back(i);
  const int nout = (int) ols.size();
  if (nout > 65535) { rhg_close(c); return fail(RHG_EUNSUPPORTED, "more than 65535 output wavelengths"); }
  {
    const int nent = m->lam_line_ptr[ns];
    std::vector<double> dv((size_t) (nent > 0 ? nent : 1));
    for (int ls = 0; ls < ns; ls++) {
      const double lambda = m->lamconst[(size_t) ls * RHG_LC_NCONST + RHG_LC_LAMBDA];
      for (int e = m->lam_line_ptr[ls]; e < m->lam_line_ptr[ls + 1]; e++)
        dv[e] = (lambda / m->line_lambda0[m->lam_line_idx[e]] - 1.0) * rhg::CLIGHT;
    }
    rc = upload(c, dv.data(), dv.size(), &d.lam_line_dv);
  }
  if (!rc) {
    const int T = rhg::BF_TILE;
    const int ntile = (nout + T - 1) / T;
    std::vector<int32_t> tptr(1, 0), tcont;
    std::vector<double> talpha;
    std::vector<int> slot(m->ncont > 0 ? m->ncont : 1);
    for (int t = 0; t < ntile; t++) {
      std::fill(slot.begin(), slot.end(), -1);
      const int base = (int) tcont.size();
      for (int j = 0; j < T && t * T + j < nout; j++) {
        const int ls = ols[t * T + j];
        for (int e = m->lam_cont_ptr[ls]; e < m->lam_cont_ptr[ls + 1]; e++) {
          const int ci = m->lam_cont_idx[e];
          if (slot[ci] < 0) {
            slot[ci] = (int) tcont.size() - base;
            tcont.push_back(ci);
            talpha.resize(talpha.size() + T, 0.0);
          }
          talpha[(size_t) (base + slot[ci]) * T + j] = m->lam_cont_alpha[e];
        }
      }
      tptr.push_back((int32_t) tcont.size());
    }
    d.ntile = ntile;
    rc = upload(c, tptr.data(), tptr.size(), &d.tile_ptr);
    if (!rc) rc = upload(c, tcont.empty() ? nullptr : tcont.data(), tcont.size(), &d.tile_cont);
    if (!rc) rc = upload(c, talpha.empty() ? nullptr : talpha.data(), talpha.size(), &d.tile_alpha);
  }
  if (!rc) {
    std::vector<double> cnt((size_t) (m->nmol > 0 ? m->nmol : 1) * rhg::NNUC_MAX, 0.0);
    for (int im = 0; im < m->nmol; im++)
      for (int j = 0; j < m->mol_nel[im]; j++) {
        const int nu = m->mol_el_nucl[4 * im + j];
        if (nu < 0 || nu >= m->nnuclei) { rc = fail(RHG_EINVAL, "molecule nucleus index out of range"); break; }
        cnt[(size_t) im * rhg::NNUC_MAX + nu] += m->mol_el_count[4 * im + j];
      }
    if (!rc) rc = upload(c, cnt.data(), cnt.size(), &d.mol_cnt);
  }
  if (rc) { std::string e = g_err; rhg_close(c); return fail(rc, e); }

  {
    for (int ls = 0; ls < ns; ls++) {
      int nc = 0;
      for (int e = m->lam_line_ptr[ls]; e < m->lam_line_ptr[ls + 1]; e++) {
        const int l = m->lam_line_idx[e];
        if (m->line_Grad[l] != 0.0 && m->line_polar[l] != 0)
          nc += m->line_comp_ptr[l + 1] - m->line_comp_ptr[l];
      }
      c->tab_maxl = std::max(c->tab_maxl, m->lam_line_ptr[ls + 1] - m->lam_line_ptr[ls]);
      c->tab_maxc = std::max(c->tab_maxc, nc);
    }
    const size_t b = rhg::linetab_bytes(std::max(c->tab_maxl, 1), std::max(c->tab_maxc, 1));
    c->tab_bytes = b <= kLineTabMaxBytes ? b : 0;
  }
  {
    std::vector<int32_t> pol, scal;
    for (int io = 0; io < nout; io++)
      (m->solver[ols[io]] == RHG_SOLVER_STOKES ? pol : scal).push_back(io);
    c->n_pol = (int) pol.size();
    c->n_scal = (int) scal.size();
    // the polarized wavelengths by the Zeeman components of their line lists,
    // most first (an order the tuning may select)
    c->h_pol = pol;
    std::vector<int> ncomp(nout, 0);
    for (int io : pol)
      for (int e = m->lam_line_ptr[ols[io]]; e < m->lam_line_ptr[ols[io] + 1]; e++) {
        const int l = m->lam_line_idx[e];
        ncomp[io] += 1 + m->line_comp_ptr[l + 1] - m->line_comp_ptr[l];
      }
    c->h_pol_heavy = pol;
    std::stable_sort(c->h_pol_heavy.begin(), c->h_pol_heavy.end(),
                     [&](int a, int b) { return ncomp[a] > ncomp[b]; });
    rc = upload(c, pol.empty() ? nullptr : pol.data(), pol.size(), &c->d_io_pol);
    if (!rc) rc = upload(c, scal.empty() ? nullptr : scal.data(), scal.size(), &c->d_io_scal);
    if (rc) { std::string e = g_err; rhg_close(c); return fail(rc, e); }
  }
  if (upload(c, &c->dm, 1, &c->d_model)) { std::string e = g_err; rhg_close(c); return fail(RHG_ENOMEM, e); }

  c->h_atom_level_ptr.assign(m->atom_level_ptr, m->atom_level_ptr + m->natom + 1);
  c->h_nucl_atom.assign(m->nucl_atom, m->nucl_atom + m->nnuclei);
  c->ref_has_line = m->lam_line_ptr[m->ref_index + 1] > m->lam_line_ptr[m->ref_index];
  c->h_level_stage.assign(m->level_stage, m->level_stage + m->atom_level_ptr[m->natom]);
  c->nspect = ns;
  c->nout = nout;
  c->ncont = m->ncont;
  c->nline = m->nline;
  c->natom = m->natom;
  if (hipMalloc(&c->d_out_ls, std::max<size_t>(ols.size(), 1) * sizeof(int32_t)) != hipSuccess) {
    rhg_close(c);
    return fail(RHG_ENOMEM, "hipMalloc (wavelength index)");
  }
  if (!ols.empty() &&
      hipMemcpy(c->d_out_ls, ols.data(), ols.size() * sizeof(int32_t), hipMemcpyHostToDevice) !=
      hipSuccess) {
    rhg_close(c);
    return fail(RHG_EDEVICE, "hipMemcpy (wavelength index)");
  }
  if (int e = noalloc_scratch_init(device)) {
    std::string msg = g_err;
    rhg_close(c);
    return fail(e, msg);
  }
  *out = c;
  g_err.clear();
  return 0;
}

// Stage 1 (k_state) for a batch: one thread per (column, depth), a
// workgroup per (column block, depth) in the XCD-aware order, the
// chemical-equilibrium kernel instantiated for the model's nucleus count.
// perm (or nullptr): slot i of the batch reads the atmosphere of column perm[i].
static int launch_state(rhg_context_t *c, const double *d_atm, const int32_t *d_ndep,
                        const int32_t *perm, int ncol, int Nmax, double mu, int atm_scale,
                        double *state, int nfield) {
  dim3 blk(64), grd((unsigned) rhg::xcd_grid((ncol + 63) / 64, Nmax));
  const size_t lds = (size_t) 2 * c->dm.nmol * 64 * sizeof(double);
  switch (c->dm.nnuclei) {
#define RHG_LAUNCH_STATE(NN)                                                                    \
  case NN:                                                                                      \
    hipLaunchKernelGGL(rhg::k_state<NN>, grd, blk, lds, c->stream, c->d_model, d_atm, d_ndep, perm, \
                       ncol, Nmax, mu, atm_scale, state, nfield);                               \
    break;
    RHG_LAUNCH_STATE(1) RHG_LAUNCH_STATE(2) RHG_LAUNCH_STATE(3) RHG_LAUNCH_STATE(4)
    RHG_LAUNCH_STATE(5) RHG_LAUNCH_STATE(6) RHG_LAUNCH_STATE(7) RHG_LAUNCH_STATE(8)
#undef RHG_LAUNCH_STATE
    default: return fail(RHG_EUNSUPPORTED, "number of molecule nuclei out of range");
  }
  HIPCHECK(hipGetLastError());
  return 0;
}

// The wavelength grids of a synthesis launch: the batch's block -> grid
// table (ncol / 64 entries, -1: a block without columns) and the context's
// per-grid wavelength lists (rhg_set_grids).
struct GridLaunch {
  const int32_t *blk_grid;
  const int32_t *pol, *pol_ptr, *scal, *scal_ptr, *tile, *mask;
  int max_pol, max_scal, mask_words;
};

// Stages 1-4 of a synthesis (k_state, k_geometry, k_continuum, the formal
// kernels) on device buffers, with the stage events; no allocation, no copy,
// no synchronisation.  d_ndep is in slot order; perm (or nullptr) maps slot
// i to the caller's column perm[i] of d_atm and d_out (-1 - c: a copy of
// column c that writes nothing).  gl (or nullptr): the batch's columns
// are grouped by wavelength grid in whole blocks of 64 slots, each block
// solving its grid's wavelengths only.
static int launch_synth(rhg_context_t *c, const double *d_atm, const int32_t *d_ndep,
                        const int32_t *perm, int ncol, int Nmax, double mu, int atm_scale,
                        int32_t precision, int nfield, double *state, int32_t *moving,
                        double *chiref, void *cchi, void *ceta, double *d_out,
                        const GridLaunch *gl = nullptr) {
  const int64_t npts = (int64_t) ncol * Nmax;
  const int nout = c->nout;

  HIPCHECK(hipEventRecord(c->ev[0], c->stream));
  {
    const int rc = launch_state(c, d_atm, d_ndep, perm, ncol, Nmax, mu, atm_scale, state, nfield);
    if (rc) return rc;
  }
  HIPCHECK(hipEventRecord(c->ev[1], c->stream));
  {
    dim3 blk(64), grd((unsigned) ((ncol + 63) / 64));
    hipLaunchKernelGGL(rhg::k_geometry, grd, blk, 0, c->stream, c->d_model, d_ndep, ncol, Nmax,
                       atm_scale, mu, state, nfield, moving, chiref);
    HIPCHECK(hipGetLastError());
  }
  HIPCHECK(hipEventRecord(c->ev[2], c->stream));
  {
    dim3 blk(rhg::CONT_BLOCK), grd((unsigned) ((npts + rhg::CONT_BLOCK - 1) / rhg::CONT_BLOCK));
    const int32_t *bg = gl ? gl->blk_grid : nullptr, *gt = gl ? gl->tile : nullptr;
    const int32_t *gm = gl ? gl->mask : nullptr;
    const int gw = gl ? gl->mask_words : 0;
#define RHG_LAUNCH_CONT(R, G)                                                                    \
  hipLaunchKernelGGL((rhg::k_continuum<R, G>), grd, blk, 0, c->stream, c->d_model, d_ndep, ncol, Nmax, \
                     state, nfield, c->d_out_ls, nout, (R *) cchi, (R *) ceta, bg, gt, gm, gw)
    if (precision == RHG_PREC_FP32) {
      if (gl) RHG_LAUNCH_CONT(float, true); else RHG_LAUNCH_CONT(float, false);
    } else {
      if (gl) RHG_LAUNCH_CONT(double, true); else RHG_LAUNCH_CONT(double, false);
    }
#undef RHG_LAUNCH_CONT
    HIPCHECK(hipGetLastError());
  }
  HIPCHECK(hipEventRecord(c->ev[3], c->stream));
  {
    const int ncb = (ncol + 63) / 64;
    dim3 blk(64);
    const int32_t *bg = gl ? gl->blk_grid : nullptr;
    const int32_t *l_pol = gl ? gl->pol : c->d_io_pol, *p_pol = gl ? gl->pol_ptr : nullptr;
    const int32_t *l_scal = gl ? gl->scal : c->d_io_scal, *p_scal = gl ? gl->scal_ptr : nullptr;
    const int n_pol = gl ? gl->max_pol : c->n_pol, n_scal = gl ? gl->max_scal : c->n_scal;
#define RHG_LAUNCH_FORMAL(K32, K64, LIST, PTR, NLIST, LDS, ...)                                  \
  if (NLIST > 0) {                                                                               \
    dim3 grd((unsigned) rhg::xcd_grid(ncb, NLIST));                                              \
    if (precision == RHG_PREC_FP32)                                                              \
      hipLaunchKernelGGL(K32, grd, blk, LDS, c->stream, c->d_model, d_ndep, moving, ncol, Nmax,  \
                         state, c->d_out_ls, LIST, NLIST, nout, (const float *) cchi,           \
                         (const float *) ceta, mu, d_out, perm, bg, PTR __VA_ARGS__);            \
    else                                                                                         \
      hipLaunchKernelGGL(K64, grd, blk, LDS, c->stream, c->d_model, d_ndep, moving, ncol, Nmax,  \
                         state, c->d_out_ls, LIST, NLIST, nout, (const double *) cchi,          \
                         (const double *) ceta, mu, d_out, perm, bg, PTR __VA_ARGS__);           \
    HIPCHECK(hipGetLastError());                                                                 \
  }
#define RHG_STOKES(MO, LT)                                                                       \
  RHG_LAUNCH_FORMAL((rhg::k_formal_stokes<float, MO, LT>), (rhg::k_formal_stokes<double, MO, LT>), \
                    l_pol, p_pol, n_pol, LT ? (unsigned) c->tab_bytes : 0u, , c->tab_maxl,       \
                    c->tab_maxc)
    const bool lt = c->tab_bytes > 0 && c->tune_line_table;
    if (c->dm.magneto_optical) {
      if (lt) { RHG_STOKES(true, true) } else { RHG_STOKES(true, false) }
    } else {
      if (lt) { RHG_STOKES(false, true) } else { RHG_STOKES(false, false) }
    }
    RHG_LAUNCH_FORMAL(rhg::k_formal_scalar<float>, rhg::k_formal_scalar<double>, l_scal, p_scal,
                      n_scal, 0u, )
#undef RHG_STOKES
#undef RHG_LAUNCH_FORMAL
  }
  HIPCHECK(hipEventRecord(c->ev[4], c->stream));
  return 0;
}

}  // extern "C"

namespace rhg {

// flag[0] = 0 (the error flag must start at zero on every call)
__global__ void k_clear_flag(int32_t *flag) {
  if (threadIdx.x == 0) flag[0] = 0;
}

// out[i] = ndep[i] clamped to [4, Nmax]; flag[0] = 1 if any was outside
__global__ void k_check_ndep(const int32_t *__restrict__ ndep, int ncol, int Nmax,
                             int32_t *__restrict__ out, int32_t *__restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncol) return;
  const int n = ndep[i];
  const bool bad = n < 4 || n > Nmax;
  out[i] = bad ? (n < 4 ? 4 : Nmax) : n;
  if (bad) atomicOr(flag, 1);
}

// The depth-count sort of a batch, on the device: a stable counting sort of
// the columns by depth count, largest first, over the values [4, Nmax]
// (out-of-range counts are clamped and counted as errors).  SORT_CHUNK
// columns per workgroup; hist is [nchunk][Nmax + 1] followed by nchunk
// out-of-range counts.
constexpr int SORT_CHUNK = 1024;

__device__ inline int clamp_ndep(int n, int Nmax) { return n < 4 ? 4 : (n > Nmax ? Nmax : n); }

// hist[g][b] = columns of chunk g whose (clamped) depth count is b;
// hist[nchunk * (Nmax + 1) + g] = columns of chunk g out of range
__global__ void __launch_bounds__(SORT_CHUNK)
k_sort_count(const int32_t *__restrict__ ndep, int ncol, int Nmax, int32_t *__restrict__ hist) {
  extern __shared__ int32_t h[];   // Nmax + 2
  const int g = blockIdx.x, nb = Nmax + 1;
  for (int b = threadIdx.x; b <= nb; b += blockDim.x) h[b] = 0;
  __syncthreads();
  const int i = g * SORT_CHUNK + threadIdx.x;
  if (i < ncol) {
    const int n = ndep[i];
    atomicAdd(&h[clamp_ndep(n, Nmax)], 1);
    if (n < 4 || n > Nmax) atomicAdd(&h[nb], 1);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += blockDim.x) hist[(int64_t) g * nb + b] = h[b];
  if (threadIdx.x == 0) hist[(int64_t) gridDim.x * nb + g] = h[nb];
}

// One workgroup: hist[g][b] becomes the first sorted position of chunk g's
// columns of depth count b (the columns of larger depth count, then those of
// count b in the chunks before g); flag[0] = the number of columns out of
// range.
__global__ void __launch_bounds__(1024)
k_sort_scan(int32_t *__restrict__ hist, int nchunk, int Nmax, int32_t *__restrict__ flag) {
  extern __shared__ int32_t tot[];   // Nmax + 2
  const int nb = Nmax + 1;
  if (threadIdx.x == 0) tot[nb] = 0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    int s = 0;
    for (int g = 0; g < nchunk; g++) s += hist[(int64_t) g * nb + b];
    tot[b] = s;
  }
  __syncthreads();
  int bad = 0;
  for (int g = threadIdx.x; g < nchunk; g += blockDim.x) bad += hist[(int64_t) nchunk * nb + g];
  if (bad) atomicAdd(&tot[nb], bad);
  if (threadIdx.x == 0) {
    int s = 0;
    for (int b = Nmax; b >= 0; b--) {
      const int t = tot[b];
      tot[b] = s;
      s += t;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) flag[0] = tot[nb];
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    int off = tot[b];
    for (int g = 0; g < nchunk; g++) {
      const int t = hist[(int64_t) g * nb + b];
      hist[(int64_t) g * nb + b] = off;
      off += t;
    }
  }
}

// Column i goes to sorted position pos = (its chunk's first position for
// its depth count) + (columns of the same count before it in the chunk):
// perm[pos] = i, sndep[pos] = its clamped depth count.
__global__ void __launch_bounds__(SORT_CHUNK)
k_sort_place(const int32_t *__restrict__ ndep, int ncol, int Nmax, const int32_t *__restrict__ off,
             int32_t *__restrict__ perm, int32_t *__restrict__ sndep) {
  __shared__ int32_t v[SORT_CHUNK];
  const int g = blockIdx.x, i = g * SORT_CHUNK + threadIdx.x;
  const int n = i < ncol ? clamp_ndep(ndep[i], Nmax) : -1;
  v[threadIdx.x] = n;
  __syncthreads();
  if (i >= ncol) return;
  int r = 0;
  for (int j = 0; j < (int) threadIdx.x; j++) r += v[j] == n;
  const int pos = off[(int64_t) g * (Nmax + 1) + n] + r;
  perm[pos] = i;
  sndep[pos] = n;
}

// The sort of a batch over wavelength grids: key (grid g, depth count n),
// the grids in order, each grid's columns largest depth count first and each
// grid's run of columns starting at a multiple of 64 slots, so that every
// block of 64 slots holds one grid.  Bins b = g * (Nmax + 1) + n; the
// histogram is k_sort_count's with ngrid * (Nmax + 1) bins per chunk.  A
// column whose grid is outside [0, ngrid) or whose depth count is out of
// range counts as an error (and is sorted as grid 0).
__device__ inline int grid_key(const int32_t *ndep, const int32_t *grid, int i, int Nmax, int ngrid,
                               bool &bad) {
  const int n = ndep[i], g = grid[i];
  const bool gok = g >= 0 && g < ngrid;
  bad = !gok || n < 4 || n > Nmax;
  return (gok ? g : 0) * (Nmax + 1) + clamp_ndep(n, Nmax);
}

__global__ void __launch_bounds__(SORT_CHUNK)
k_gsort_count(const int32_t *__restrict__ ndep, const int32_t *__restrict__ grid, int ncol, int Nmax,
              int ngrid, int32_t *__restrict__ hist) {
  extern __shared__ int32_t h[];   // nb + 1
  const int g = blockIdx.x, nb = ngrid * (Nmax + 1);
  for (int b = threadIdx.x; b <= nb; b += blockDim.x) h[b] = 0;
  __syncthreads();
  const int i = g * SORT_CHUNK + threadIdx.x;
  if (i < ncol) {
    bool bad;
    atomicAdd(&h[grid_key(ndep, grid, i, Nmax, ngrid, bad)], 1);
    if (bad) atomicAdd(&h[nb], 1);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += blockDim.x) hist[(int64_t) g * nb + b] = h[b];
  if (threadIdx.x == 0) hist[(int64_t) gridDim.x * nb + g] = h[nb];
}

// One workgroup: hist[chunk][b] becomes the first sorted slot of the chunk's
// columns of bin b; seg[3 g .. 3 g + 2] = grid g's first slot, the end of its
// columns and the end of its run padded to a multiple of 64; flag[0] = the
// number of columns in error.
__global__ void __launch_bounds__(1024)
k_gsort_scan(int32_t *__restrict__ hist, int nchunk, int Nmax, int ngrid, int32_t *__restrict__ seg,
             int32_t *__restrict__ flag) {
  extern __shared__ int32_t tot[];   // nb + 1
  const int nb = ngrid * (Nmax + 1);
  if (threadIdx.x == 0) tot[nb] = 0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    int s = 0;
    for (int g = 0; g < nchunk; g++) s += hist[(int64_t) g * nb + b];
    tot[b] = s;
  }
  __syncthreads();
  int bad = 0;
  for (int g = threadIdx.x; g < nchunk; g += blockDim.x) bad += hist[(int64_t) nchunk * nb + g];
  if (bad) atomicAdd(&tot[nb], bad);
  if (threadIdx.x == 0) {
    int s = 0;
    for (int q = 0; q < ngrid; q++) {
      s = (s + 63) & ~63;
      seg[3 * q] = s;
      for (int n = Nmax; n >= 0; n--) {
        const int b = q * (Nmax + 1) + n, t = tot[b];
        tot[b] = s;
        s += t;
      }
      seg[3 * q + 1] = s;
      seg[3 * q + 2] = (s + 63) & ~63;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) flag[0] = tot[nb];
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    int off = tot[b];
    for (int g = 0; g < nchunk; g++) {
      const int t = hist[(int64_t) g * nb + b];
      hist[(int64_t) g * nb + b] = off;
      off += t;
    }
  }
}

// Column i goes to slot (its chunk's first slot for its bin) + (columns of
// the same bin before it in the chunk): perm[slot] = i, sndep[slot] = its
// clamped depth count.
__global__ void __launch_bounds__(SORT_CHUNK)
k_gsort_place(const int32_t *__restrict__ ndep, const int32_t *__restrict__ grid, int ncol, int Nmax,
              int ngrid, const int32_t *__restrict__ off, int32_t *__restrict__ perm,
              int32_t *__restrict__ sndep) {
  __shared__ int32_t v[SORT_CHUNK];
  const int g = blockIdx.x, i = g * SORT_CHUNK + threadIdx.x;
  bool bad = false;
  const int key = i < ncol ? grid_key(ndep, grid, i, Nmax, ngrid, bad) : -1;
  v[threadIdx.x] = key;
  __syncthreads();
  if (i >= ncol) return;
  int r = 0;
  for (int j = 0; j < (int) threadIdx.x; j++) r += v[j] == key;
  const int pos = off[(int64_t) g * ngrid * (Nmax + 1) + key] + r;
  perm[pos] = i;
  sndep[pos] = clamp_ndep(ndep[i], Nmax);
}

// Every slot of [0, nslot) not holding a column: a slot that pads grid g's
// run to whole blocks becomes a copy of the grid's last column (perm entry
// -1 - column: read, never written), a slot past every run a copy of column 0
// with 4 depths; blk[slot / 64] = the grid of the block (-1: none).
__global__ void k_gsort_pad(const int32_t *__restrict__ seg, int ngrid, int nslot,
                            int32_t *__restrict__ perm, int32_t *__restrict__ sndep,
                            int32_t *__restrict__ blk) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslot) return;
  int g = -1;
  for (int q = 0; q < ngrid; q++)
    if (s >= seg[3 * q] && s < seg[3 * q + 2]) g = q;
  if ((s & 63) == 0) blk[s >> 6] = g;
  if (g < 0) {
    perm[s] = -1;
    sndep[s] = 4;
  } else if (s >= seg[3 * g + 1]) {
    const int src = seg[3 * g + 1] - 1;
    perm[s] = -1 - perm[src];
    sndep[s] = sndep[src];
  }
}

}  // namespace rhg

namespace {

// The depth counts the synthesis kernels read, on the device, without a
// copy to the host: the caller's d_ndep checked and clamped to [4, Nmax]
// into syn_ndep, and for a batch of 128 columns or more (column sorting on)
// sorted by depth count, largest first, with syn_perm[slot] = the caller's
// column (a wavefront of 64 columns waits for its deepest column).  The
// flag (syn_hist's last word) ends as the number of columns out of range.
// Sets *sndep_out, *perm_out (nullptr: slot = column) and *flag_out.
int prepare_ndep(rhg_context_t *c, const int32_t *d_ndep, int ncol, int Nmax, bool sort,
                 const int32_t **sndep_out, const int32_t **perm_out, int32_t **flag_out) {
  sort = sort && c->tune_column_sort && ncol >= 128;
  const int nchunk = (ncol + rhg::SORT_CHUNK - 1) / rhg::SORT_CHUNK;
  const size_t hist_n = sort ? (size_t) nchunk * (Nmax + 2) : 0;
  if (c->syn_ndep.ensure((size_t) ncol * sizeof(int32_t)) ||
      (sort && c->syn_perm.ensure((size_t) ncol * sizeof(int32_t))) ||
      c->syn_hist.ensure((hist_n + 1) * sizeof(int32_t)))
    return fail(RHG_ENOMEM, "device allocation for the batch failed");
  int32_t *sndep = (int32_t *) c->syn_ndep.p, *hist = (int32_t *) c->syn_hist.p;
  int32_t *flag = hist + hist_n;
  const size_t lds = (size_t) (Nmax + 2) * sizeof(int32_t);
  if (sort) {
    int32_t *perm = (int32_t *) c->syn_perm.p;
    hipLaunchKernelGGL(rhg::k_sort_count, dim3((unsigned) nchunk), dim3(rhg::SORT_CHUNK), lds,
                       c->stream, d_ndep, ncol, Nmax, hist);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(rhg::k_sort_scan, dim3(1), dim3(1024), lds, c->stream, hist, nchunk, Nmax, flag);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(rhg::k_sort_place, dim3((unsigned) nchunk), dim3(rhg::SORT_CHUNK), 0,
                       c->stream, d_ndep, ncol, Nmax, hist, perm, sndep);
    HIPCHECK(hipGetLastError());
    *perm_out = perm;
  } else {
    hipLaunchKernelGGL(rhg::k_clear_flag, dim3(1), dim3(64), 0, c->stream, flag);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(rhg::k_check_ndep, dim3((unsigned) ((ncol + 255) / 256)), dim3(256), 0,
                       c->stream, d_ndep, ncol, Nmax, sndep, flag);
    HIPCHECK(hipGetLastError());
    *perm_out = nullptr;
  }
  *sndep_out = sndep;
  *flag_out = flag;
  return 0;
}

// The wavelength-grid form of prepare_ndep: the caller's columns (grid ids
// d_grid) sorted by grid and depth count into nslot = 64 (ceil(ncol / 64) +
// ngrid) slots (every grid's run a whole number of blocks; the block ->
// grid table in gr_blk); the flag ends as the number of columns whose depth
// count or grid is out of range.
int prepare_grids(rhg_context_t *c, const int32_t *d_ndep, const int32_t *d_grid, int ncol,
                  int Nmax, int nslot, const int32_t **sndep_out, const int32_t **perm_out,
                  int32_t **flag_out) {
  const int ngrid = c->ngrid, nb = ngrid * (Nmax + 1);
  const int nchunk = (ncol + rhg::SORT_CHUNK - 1) / rhg::SORT_CHUNK;
  const size_t lds = (size_t) (nb + 1) * sizeof(int32_t);
  if (lds > 65536) return fail(RHG_EINVAL, "too many wavelength grids for this ndep_max");
  const size_t hist_n = (size_t) nchunk * (nb + 1);
  if (c->syn_ndep.ensure((size_t) nslot * sizeof(int32_t)) ||
      c->syn_perm.ensure((size_t) nslot * sizeof(int32_t)) ||
      c->syn_hist.ensure((hist_n + 3 * (size_t) ngrid + 1) * sizeof(int32_t)) ||
      c->gr_blk.ensure((size_t) (nslot / 64) * sizeof(int32_t)))
    return fail(RHG_ENOMEM, "device allocation for the batch failed");
  int32_t *sndep = (int32_t *) c->syn_ndep.p, *perm = (int32_t *) c->syn_perm.p;
  int32_t *hist = (int32_t *) c->syn_hist.p, *seg = hist + hist_n, *flag = seg + 3 * ngrid;
  hipLaunchKernelGGL(rhg::k_gsort_count, dim3((unsigned) nchunk), dim3(rhg::SORT_CHUNK), lds,
                     c->stream, d_ndep, d_grid, ncol, Nmax, ngrid, hist);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(rhg::k_gsort_scan, dim3(1), dim3(1024), lds, c->stream, hist, nchunk, Nmax,
                     ngrid, seg, flag);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(rhg::k_gsort_place, dim3((unsigned) nchunk), dim3(rhg::SORT_CHUNK), 0,
                     c->stream, d_ndep, d_grid, ncol, Nmax, ngrid, hist, perm, sndep);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(rhg::k_gsort_pad, dim3((unsigned) ((nslot + 255) / 256)), dim3(256), 0,
                     c->stream, seg, ngrid, nslot, perm, sndep, (int32_t *) c->gr_blk.p);
  HIPCHECK(hipGetLastError());
  *sndep_out = sndep;
  *perm_out = perm;
  *flag_out = flag;
  return 0;
}

// The synthesis of a batch.  Host buffers (device_ptrs 0) are checked on
// the host and copied in; device buffers are checked on the device (no copy
// to the host before the launches).  sort: run a batch of uneven depth
// counts in depth-count order (the kernels read the caller's atmosphere and
// write its Stokes vectors through the permutation); the diagnostic outputs
// and the callers that reuse the context's column state take the batch in
// the caller's order.  col_grid (or nullptr): the wavelength grid of every
// column (rhg_set_grids), host or device memory as the other inputs; the
// batch then runs grouped by grid, each column solved at its grid's
// wavelengths only, the other entries of its Stokes output set to zero.
int synth(rhg_context_t *c, const rhg_columns_t *cols, int32_t precision, int32_t device_ptrs,
          bool sort, double *stokes_out, double *height_out, double *chiref_out,
          const int32_t *col_grid = nullptr) {
  if (!c || !cols || !stokes_out || !cols->atmos || !cols->ndep) return fail(RHG_EINVAL, "null argument");
  if (c->nout < 1) return fail(RHG_EINVAL, "the model has no output wavelength");
  if (precision != RHG_PREC_FP64 && precision != RHG_PREC_FP32)
    return fail(RHG_EINVAL, "precision must be RHG_PREC_FP64 or RHG_PREC_FP32");
  const int ncol = cols->ncol, Nmax = cols->ndep_max, nout = c->nout;
  if (ncol < 1 || Nmax < 4 || Nmax > RHG_MAX_DEPTH) return fail(RHG_EINVAL, "bad batch shape");
  if ((int64_t) ncol * Nmax > (int64_t) INT32_MAX) return fail(RHG_EINVAL, "batch too large");
  if (!(cols->mu > 0.0 && cols->mu <= 1.0)) return fail(RHG_EINVAL, "mu must be in (0, 1]");
  if (cols->atm_scale < 0 || cols->atm_scale > 2) return fail(RHG_EINVAL, "bad atm_scale");
  if (device_ptrs && height_out) return fail(RHG_EINVAL, "height output needs host buffers");
  if (!device_ptrs)
    for (int i = 0; i < ncol; i++)
      if (cols->ndep[i] < 4 || cols->ndep[i] > Nmax) return fail(RHG_EINVAL, "ndep out of range");
  if (col_grid) {
    if (c->ngrid < 1) return fail(RHG_EINVAL, "no wavelength grids are set (rhg_set_grids)");
    if (height_out || chiref_out) return fail(RHG_EINVAL, "grid batches have no diagnostic outputs");
    if (!device_ptrs)
      for (int i = 0; i < ncol; i++)
        if (col_grid[i] < 0 || col_grid[i] >= c->ngrid) return fail(RHG_EINVAL, "grid out of range");
  }
  HIPCHECK(hipSetDevice(c->device));
  if (height_out || chiref_out) sort = false;
  // slots of the kernels: the columns, or with grids every grid's columns
  // padded to whole blocks of 64
  const int nslot = col_grid ? 64 * ((ncol + 63) / 64 + c->ngrid) : ncol;
  if ((int64_t) nslot * Nmax > (int64_t) INT32_MAX) return fail(RHG_EINVAL, "batch too large");

  const int nfield = rhg::state_nfield(c->ncont, c->nline);
  const size_t rsz = precision == RHG_PREC_FP32 ? sizeof(float) : sizeof(double);
  const size_t atm_b = (size_t) ncol * 9 * Nmax * sizeof(double);
  const size_t st_b = (size_t) nslot * nfield * Nmax * sizeof(double);
  const size_t cc_b = (size_t) nslot * Nmax * nout * rsz;
  const size_t out_b = (size_t) ncol * 4 * nout * sizeof(double);
  if (c->state.ensure(st_b) || c->moving.ensure(nslot * sizeof(int32_t)) ||
      c->cchi.ensure(cc_b) || c->ceta.ensure(cc_b) ||
      (!device_ptrs && (c->atmos.ensure(atm_b) || c->ndep.ensure(ncol * sizeof(int32_t)) ||
                        c->stokes.ensure(out_b))) ||
      (col_grid && !device_ptrs && c->gr_col.ensure(ncol * sizeof(int32_t))) ||
      (chiref_out && c->chiref.ensure((size_t) ncol * Nmax * sizeof(double))))
    return fail(RHG_ENOMEM, "device allocation for the batch failed");
  const double *d_atm = (const double *) cols->atmos;
  const int32_t *d_ndep = cols->ndep;
  double *d_out = stokes_out;
  if (!device_ptrs) {
    HIPCHECK(hipMemcpyAsync(c->atmos.p, cols->atmos, atm_b, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->ndep.p, cols->ndep, ncol * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    d_atm = (const double *) c->atmos.p;
    d_ndep = (const int32_t *) c->ndep.p;
    d_out = (double *) c->stokes.p;
  }
  const int32_t *sndep = nullptr, *perm = nullptr;
  int32_t *flag = nullptr;
  GridLaunch gl{};
  if (col_grid) {
    const int32_t *d_grid = col_grid;
    if (!device_ptrs) {
      HIPCHECK(hipMemcpyAsync(c->gr_col.p, col_grid, ncol * sizeof(int32_t), hipMemcpyHostToDevice,
                              c->stream));
      d_grid = (const int32_t *) c->gr_col.p;
    }
    if (int rc = prepare_grids(c, d_ndep, d_grid, ncol, Nmax, nslot, &sndep, &perm, &flag)) return rc;
    gl = GridLaunch{(const int32_t *) c->gr_blk.p, (const int32_t *) c->gr_pol.p,
                    (const int32_t *) c->gr_pol_ptr.p, (const int32_t *) c->gr_scal.p,
                    (const int32_t *) c->gr_scal_ptr.p, (const int32_t *) c->gr_tile.p,
                    (const int32_t *) c->gr_mask.p, c->gr_max_pol, c->gr_max_scal, c->gr_words};
    HIPCHECK(hipMemsetAsync(d_out, 0, out_b, c->stream));
  } else if (int rc = prepare_ndep(c, d_ndep, ncol, Nmax, sort, &sndep, &perm, &flag)) {
    return rc;
  }
  double *state = (double *) c->state.p;
  int32_t *moving = (int32_t *) c->moving.p;
  double *chiref = chiref_out ? (double *) c->chiref.p : nullptr;
  if (int rc = launch_synth(c, d_atm, sndep, perm, nslot, Nmax, cols->mu, cols->atm_scale,
                            precision, nfield, state, moving, chiref, c->cchi.p, c->ceta.p, d_out,
                            col_grid ? &gl : nullptr))
    return rc;
  if (!device_ptrs)
    HIPCHECK(hipMemcpyAsync(stokes_out, d_out, out_b, hipMemcpyDeviceToHost, c->stream));
  std::vector<double> hfield;
  if (height_out) {
    hfield.resize((size_t) Nmax * ncol);
    HIPCHECK(hipMemcpyAsync(hfield.data(), state + (size_t) RHG_CF_HEIGHT * Nmax * ncol,
                            hfield.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  if (chiref_out)
    HIPCHECK(hipMemcpyAsync(chiref_out, chiref, (size_t) ncol * Nmax * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
  int32_t nbad = 0;
  HIPCHECK(hipMemcpyAsync(&nbad, flag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (nbad) return fail(RHG_EINVAL, col_grid ? "ndep or grid out of range" : "ndep out of range");
  if (height_out)
    for (int ic = 0; ic < ncol; ic++)
      for (int k = 0; k < Nmax; k++)
        height_out[(size_t) ic * Nmax + k] =
            k < cols->ndep[ic] ? hfield[(size_t) k * ncol + ic] : 0.0;
  // the work census reads the last batch's slots in the caller's order
  // (or sorted by depth count); a grid batch's slots are not that
  c->last_ncol = col_grid ? 0 : ncol;
  c->last_Nmax = Nmax;
  c->last_nfield = nfield;
  c->last_mu = cols->mu;
  c->last_ndep = sndep;
  // events of the stages: state, geometry, continuum, formal
  float ms;
  for (int s = 0; s < 4; s++) {
    HIPCHECK(hipEventElapsedTime(&ms, c->ev[s], c->ev[s + 1]));
    c->last_ms[s] = ms;
  }
  g_err.clear();
  return 0;
}

}  // namespace



namespace {

// A response-function batch whose depth counts differ runs with its columns
// gathered in depth-count order, longest first (the synthesis sorts on the
// device instead, prepare_ndep): a wavefront holds 64 columns at one
// wavelength, and each lane's depth recursion is as long as its column, so a
// wavefront of mixed depths waits for its deepest column.  Returns the
// permutation (empty when the batch is small or already ordered); perm[i] is
// the caller's column at sorted position i.
int sort_order(rhg_context_t *c, const rhg_columns_t *cols, int32_t device_ptrs,
               std::vector<int32_t> &perm, std::vector<int32_t> &sorted_ndep) {
  perm.clear();
  const int ncol = cols->ncol;
  if (ncol < 128 || !cols->ndep || !c->tune_column_sort) return 0;
  std::vector<int32_t> hndep(ncol);
  if (device_ptrs) {
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipMemcpy(hndep.data(), cols->ndep, ncol * sizeof(int32_t), hipMemcpyDeviceToHost));
  } else {
    memcpy(hndep.data(), cols->ndep, ncol * sizeof(int32_t));
  }
  bool ordered = true;
  for (int i = 1; i < ncol; i++) ordered &= hndep[i] <= hndep[i - 1];
  if (ordered) return 0;
  perm.resize(ncol);
  for (int i = 0; i < ncol; i++) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return hndep[a] > hndep[b]; });
  sorted_ndep.resize(ncol);
  for (int i = 0; i < ncol; i++) sorted_ndep[i] = hndep[perm[i]];
  return 0;
}

// Gather the caller's atmospheres into the sorted order; fills cols2.
int gather_sorted(rhg_context_t *c, const rhg_columns_t *cols, int32_t device_ptrs,
                  const std::vector<int32_t> &perm, const std::vector<int32_t> &sorted_ndep,
                  rhg_columns_t &cols2) {
  const int ncol = cols->ncol, Nmax = cols->ndep_max;
  const size_t atm_b = (size_t) ncol * 9 * Nmax * sizeof(double);
  if (c->srt_atm.ensure(atm_b) || c->srt_ndep.ensure(ncol * sizeof(int32_t)) ||
      c->srt_perm.ensure(ncol * sizeof(int32_t)) || (!device_ptrs && c->atmos.ensure(atm_b)))
    return fail(RHG_ENOMEM, "device allocation for the sorted batch failed");
  HIPCHECK(hipSetDevice(c->device));
  const double *src = (const double *) cols->atmos;
  if (!device_ptrs) {
    HIPCHECK(hipMemcpyAsync(c->atmos.p, cols->atmos, atm_b, hipMemcpyHostToDevice, c->stream));
    src = (const double *) c->atmos.p;
  }
  HIPCHECK(hipMemcpyAsync(c->srt_perm.p, perm.data(), ncol * sizeof(int32_t), hipMemcpyHostToDevice,
                          c->stream));
  HIPCHECK(hipMemcpyAsync(c->srt_ndep.p, sorted_ndep.data(), ncol * sizeof(int32_t),
                          hipMemcpyHostToDevice, c->stream));
  const int64_t n = (int64_t) ncol * 9 * Nmax;
  hipLaunchKernelGGL(rhg::k_gather_cols, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, c->stream,
                     src, (double *) c->srt_atm.p, (const int32_t *) c->srt_perm.p, ncol,
                     (int64_t) 9 * Nmax);
  HIPCHECK(hipGetLastError());
  cols2 = *cols;
  cols2.atmos = (const double *) c->srt_atm.p;
  cols2.ndep = (const int32_t *) c->srt_ndep.p;
  return 0;
}

// Stokes [ncol][4][nout] from the sorted order back to the caller's.
int scatter_stokes(rhg_context_t *c, int ncol, int32_t device_ptrs, const double *src,
                   double *stokes_out) {
  const int64_t len = (int64_t) 4 * c->nout;
  const size_t b = (size_t) ncol * len * sizeof(double);
  if (!device_ptrs && c->stokes.ensure(b)) return fail(RHG_ENOMEM, "device allocation failed");
  double *dst = device_ptrs ? stokes_out : (double *) c->stokes.p;
  const int64_t n = (int64_t) ncol * len;
  hipLaunchKernelGGL(rhg::k_scatter_cols, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, c->stream,
                     src, dst, (const int32_t *) c->srt_perm.p, ncol, len);
  HIPCHECK(hipGetLastError());
  if (!device_ptrs) HIPCHECK(hipMemcpyAsync(stokes_out, dst, b, hipMemcpyDeviceToHost, c->stream));
  return 0;
}

}  // namespace

extern "C" {

int rhg_synth_lte(rhg_context_t *ctx, const rhg_columns_t *cols, int32_t precision,
                  int32_t device_ptrs, double *stokes_out) {
  return synth(ctx, cols, precision, device_ptrs, true, stokes_out, nullptr, nullptr);
}

int rhg_synth_lte_diag(rhg_context_t *ctx, const rhg_columns_t *cols, int32_t precision,
                       double *stokes_out, double *height_out, double *chiref_out) {
  return synth(ctx, cols, precision, 0, false, stokes_out, height_out, chiref_out);
}

int rhg_set_grids(rhg_context_t *c, int32_t ngrid, const int32_t *grid_ptr,
                  const int32_t *grid_out) {
  if (!c || !grid_ptr || (!grid_out && ngrid > 0)) return fail(RHG_EINVAL, "null argument");
  if (ngrid < 1 || ngrid > RHG_MAX_GRIDS) return fail(RHG_EINVAL, "ngrid out of range");
  if (grid_ptr[0] != 0) return fail(RHG_EINVAL, "grid_ptr[0] must be 0");
  const int nout = c->nout;
  std::vector<char> is_pol(nout, 0);
  for (int io : c->h_pol) is_pol[io] = 1;
  std::vector<int32_t> pol, scal, pol_ptr{0}, scal_ptr{0}, tile;
  const int words = (nout + 31) / 32;
  std::vector<int32_t> mask((size_t) ngrid * words, 0);
  int max_pol = 0, max_scal = 0;
  for (int g = 0; g < ngrid; g++) {
    if (grid_ptr[g + 1] < grid_ptr[g]) return fail(RHG_EINVAL, "grid_ptr must not decrease");
    int lo = nout, hi = -1;
    for (int e = grid_ptr[g]; e < grid_ptr[g + 1]; e++) {
      const int io = grid_out[e];
      if (io < 0 || io >= nout) return fail(RHG_EINVAL, "grid wavelength position out of range");
      if (e > grid_ptr[g] && io <= grid_out[e - 1])
        return fail(RHG_EINVAL, "a grid's wavelength positions must increase");
      (is_pol[io] ? pol : scal).push_back(io);
      mask[(size_t) g * words + io / 32] |= (int32_t) (1u << (io % 32));
      lo = std::min(lo, io);
      hi = std::max(hi, io);
    }
    max_pol = std::max(max_pol, (int) pol.size() - pol_ptr.back());
    max_scal = std::max(max_scal, (int) scal.size() - scal_ptr.back());
    pol_ptr.push_back((int32_t) pol.size());
    scal_ptr.push_back((int32_t) scal.size());
    tile.push_back(hi < 0 ? 0 : lo / rhg::BF_TILE);
    tile.push_back(hi < 0 ? 0 : hi / rhg::BF_TILE + 1);
  }
  HIPCHECK(hipSetDevice(c->device));
  auto put = [&](DevBuf &b, const std::vector<int32_t> &v) -> int {
    const size_t n = std::max<size_t>(v.size(), 1) * sizeof(int32_t);
    if (b.ensure(n)) return fail(RHG_ENOMEM, "device allocation for the grids failed");
    if (!v.empty()) HIPCHECK(hipMemcpy(b.p, v.data(), v.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    return 0;
  };
  c->ngrid = 0;
  if (int rc = put(c->gr_pol, pol)) return rc;
  if (int rc = put(c->gr_pol_ptr, pol_ptr)) return rc;
  if (int rc = put(c->gr_scal, scal)) return rc;
  if (int rc = put(c->gr_scal_ptr, scal_ptr)) return rc;
  if (int rc = put(c->gr_tile, tile)) return rc;
  if (int rc = put(c->gr_mask, mask)) return rc;
  c->ngrid = ngrid;
  c->gr_words = words;
  c->gr_max_pol = max_pol;
  c->gr_max_scal = max_scal;
  g_err.clear();
  return 0;
}

int rhg_synth_lte_grids(rhg_context_t *ctx, const rhg_columns_t *cols, int32_t precision,
                        int32_t device_ptrs, const int32_t *col_grid, double *stokes_out) {
  if (!col_grid) return fail(RHG_EINVAL, "null argument");
  return synth(ctx, cols, precision, device_ptrs, true, stokes_out, nullptr, nullptr, col_grid);
}

}  // extern "C"

namespace {

size_t align256(size_t n) { return (n + 255) & ~(size_t) 255; }

}  // namespace

extern "C" {

int rhg_synth_lte_noalloc(rhg_context_t *ctx, const rhg_columns_t *cols, int32_t precision,
                          double *stokes_out) {
  if (!ctx || !cols || !cols->atmos || !cols->ndep || !stokes_out)
    return fail(RHG_EINVAL, "null argument");
  if (ctx->nout < 1) return fail(RHG_EINVAL, "the model has no output wavelength");
  if (precision != RHG_PREC_FP64 && precision != RHG_PREC_FP32)
    return fail(RHG_EINVAL, "precision must be RHG_PREC_FP64 or RHG_PREC_FP32");
  const int ncol = cols->ncol, Nmax = cols->ndep_max, nout = ctx->nout;
  if (ncol < 1 || Nmax < 4 || Nmax > RHG_MAX_DEPTH) return fail(RHG_EINVAL, "bad batch shape");
  if (!(cols->mu > 0.0 && cols->mu <= 1.0)) return fail(RHG_EINVAL, "mu must be in (0, 1]");
  if (cols->atm_scale < 0 || cols->atm_scale > 2) return fail(RHG_EINVAL, "bad atm_scale");
  const int dev = ctx->device;
  if (dev < 0 || dev >= kMaxDevices || !g_scratch_ptr[dev].load())
    return fail(RHG_EINVAL, "no-allocation scratch not set up on this device");
  const int nfield = rhg::state_nfield(ctx->ncont, ctx->nline);
  const size_t rsz = precision == RHG_PREC_FP32 ? sizeof(float) : sizeof(double);
  const size_t npts = (size_t) ncol * Nmax;
  const size_t b_ndep = align256((size_t) ncol * sizeof(int32_t)), b_flag = 256,
               b_mov = align256((size_t) ncol * sizeof(int32_t)),
               b_state = align256(npts * nfield * sizeof(double)),
               b_cont = align256(npts * nout * rsz);
  const size_t need = b_ndep + b_flag + b_mov + b_state + 2 * b_cont;
  if (need > (size_t) RHG_NOALLOC_SCRATCH_BYTES) {
    const size_t per_col = need / (size_t) ncol + 1;
    return fail(RHG_EINVAL, "batch needs " + std::to_string(need >> 20) + " MiB of work space, "
                                "beyond the no-allocation scratch of " +
                                std::to_string(RHG_NOALLOC_SCRATCH_BYTES >> 20) +
                                " MiB: at most about " +
                                std::to_string((size_t) RHG_NOALLOC_SCRATCH_BYTES / per_col) +
                                " columns of this shape per call");
  }
  HIPCHECK(hipSetDevice(dev));
  unsigned char *base = (unsigned char *) g_scratch_ptr[dev].load();
  int32_t *sndep = (int32_t *) base;
  int32_t *flag = (int32_t *) (base + b_ndep);
  int32_t *moving = (int32_t *) (base + b_ndep + b_flag);
  double *state = (double *) (base + b_ndep + b_flag + b_mov);
  void *cchi = base + b_ndep + b_flag + b_mov + b_state;
  void *ceta = (unsigned char *) cchi + b_cont;
  hipLaunchKernelGGL(rhg::k_clear_flag, dim3(1), dim3(64), 0, ctx->stream, flag);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(rhg::k_check_ndep, dim3((unsigned) ((ncol + 255) / 256)), dim3(256), 0,
                     ctx->stream, cols->ndep, ncol, Nmax, sndep, flag);
  HIPCHECK(hipGetLastError());
  if (int rc = launch_synth(ctx, (const double *) cols->atmos, sndep, nullptr, ncol, Nmax, cols->mu,
                            cols->atm_scale, precision, nfield, state, moving, nullptr, cchi,
                            ceta, stokes_out))
    return rc;
  int32_t hflag = 0;
  HIPCHECK(hipMemcpyAsync(&hflag, flag, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  if (hflag) return fail(RHG_EINVAL, "ndep out of range");
  g_err.clear();
  return 0;
}

int rhg_count_work(rhg_context_t *c, int64_t *counts_out) {
  if (!c || !counts_out) return fail(RHG_EINVAL, "null argument");
  if (c->last_ncol < 1) return fail(RHG_EINVAL, "no batch has been synthesised on this context");
  HIPCHECK(hipSetDevice(c->device));
  const int ncol = c->last_ncol, Nmax = c->last_Nmax, nout = c->nout, nfield = c->last_nfield;
  if (c->census.ensure((size_t) ncol * nout * RHG_NWORK * sizeof(int32_t)) ||
      c->census_sum.ensure((size_t) ncol * RHG_NWORK * sizeof(int64_t)))
    return fail(RHG_ENOMEM, "device allocation for the work census failed");
  int32_t *census = (int32_t *) c->census.p;
  const double *state = (const double *) c->state.p;
  {
    dim3 blk(64), grd((unsigned) ((ncol + 63) / 64), (unsigned) nout);
    hipLaunchKernelGGL(rhg::k_census, grd, blk, 0, c->stream, c->d_model, c->last_ndep,
                       (const int32_t *) c->moving.p, ncol, Nmax, state, nfield, c->d_out_ls, nout,
                       c->last_mu, census);
    HIPCHECK(hipGetLastError());
  }
  {
    dim3 blk(64), grd((unsigned) ((ncol + 63) / 64));
    hipLaunchKernelGGL(rhg::k_census_sum, grd, blk, 0, c->stream, census, ncol, nout,
                       (int64_t *) c->census_sum.p);
    HIPCHECK(hipGetLastError());
  }
  std::vector<int64_t> cs((size_t) ncol * RHG_NWORK);
  HIPCHECK(hipMemcpyAsync(cs.data(), c->census_sum.p, cs.size() * sizeof(int64_t),
                          hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < RHG_NWORK; i++) counts_out[i] = 0;
  for (int ic = 0; ic < ncol; ic++)
    for (int i = 0; i < RHG_NWORK; i++) counts_out[i] += cs[(size_t) ic * RHG_NWORK + i];
  std::vector<int32_t> hn(ncol);
  HIPCHECK(hipMemcpy(hn.data(), c->last_ndep, ncol * sizeof(int32_t), hipMemcpyDeviceToHost));
  int64_t cd = 0;
  for (int ic = 0; ic < ncol; ic++) cd += hn[ic];
  counts_out[RHG_WORK_COLDEPTH] = cd;
  g_err.clear();
  return 0;
}

}  // extern "C"

namespace {

// k_state + k_geometry + k_continuum<R> on one batch (no events, no copies)
template <typename R>
int launch_base(rhg_context_t *c, const double *d_atm, const int32_t *d_ndep, int ncol, int Nmax,
                double mu, int atm_scale, double *state, int nfield, int32_t *moving,
                double *chiref, R *cchi, R *ceta) {
  const int64_t npts = (int64_t) ncol * Nmax;
  {
    dim3 blk(64), grd((unsigned) rhg::xcd_grid((ncol + 63) / 64, Nmax));
    const size_t lds = (size_t) 2 * c->dm.nmol * 64 * sizeof(double);
    switch (c->dm.nnuclei) {
#define RHG_LAUNCH_STATE(NN)                                                                   \
  case NN:                                                                                     \
    hipLaunchKernelGGL(rhg::k_state<NN>, grd, blk, lds, c->stream, c->d_model, d_atm, d_ndep,     \
                       nullptr, ncol, Nmax, mu, atm_scale, state, nfield);                     \
    break;
      RHG_LAUNCH_STATE(1) RHG_LAUNCH_STATE(2) RHG_LAUNCH_STATE(3) RHG_LAUNCH_STATE(4)
      RHG_LAUNCH_STATE(5) RHG_LAUNCH_STATE(6) RHG_LAUNCH_STATE(7) RHG_LAUNCH_STATE(8)
#undef RHG_LAUNCH_STATE
      default: return fail(RHG_EUNSUPPORTED, "number of molecule nuclei out of range");
    }
    HIPCHECK(hipGetLastError());
  }
  {
    dim3 blk(64), grd((unsigned) ((ncol + 63) / 64));
    hipLaunchKernelGGL(rhg::k_geometry, grd, blk, 0, c->stream, c->d_model, d_ndep, ncol, Nmax,
                       atm_scale, mu, state, nfield, moving, chiref);
    HIPCHECK(hipGetLastError());
  }
  if (cchi) {
    dim3 blk(rhg::CONT_BLOCK), grd((unsigned) ((npts + rhg::CONT_BLOCK - 1) / rhg::CONT_BLOCK));
    hipLaunchKernelGGL((rhg::k_continuum<R, false>), grd, blk, 0, c->stream, c->d_model, d_ndep, ncol, Nmax,
                       state, nfield, c->d_out_ls, c->nout, cchi, ceta, (const int32_t *) nullptr, (const int32_t *) nullptr, (const int32_t *) nullptr, 0);
    HIPCHECK(hipGetLastError());
  }
  return 0;
}

// the state tangent steps (atmosphere row, step in the row's units)
struct RfParam { int row; double h; };
const RfParam kRfParam[RHG_RF_NPARAM] = {
    {1, 1.0e-2},   // T [K]
    {3, 1.0e-4},   // v_LOS [km/s]
    {5, 1.0e-1},   // B [G]
    {6, 1.0e-6},   // gamma [rad]
    {7, 1.0e-6},   // chi [rad]
    {4, 1.0e-4},   // v_mic [km/s]
};

template <typename R>
int rf_chunk(rhg_context_t *c, const double *d_atm, const int32_t *d_ndep, int ncol, int Nmax,
             double mu, int scale, const rhg::RfParamList &pl, double *d_out, R *d_rf,
             int64_t ldr) {
  const int nout = c->nout;
  const int nfield = rhg::state_nfield(c->ncont, c->nline);
  const int64_t npts = (int64_t) ncol * Nmax;
  const int npol = c->n_pol;
  double *state = (double *) c->state.p;
  int32_t *moving = (int32_t *) c->moving.p;
  R *cchi = (R *) c->cchi.p, *ceta = (R *) c->ceta.p;
  double *chiref = (double *) c->chiref.p;
  HIPCHECK(hipEventRecord(c->ev[0], c->stream));
  int rc = launch_base<R>(c, d_atm, d_ndep, ncol, Nmax, mu, scale, state, nfield, moving, chiref,
                          cchi, ceta);
  if (rc) return rc;
  HIPCHECK(hipEventRecord(c->ev[1], c->stream));

  // 1. state tangents by central differences (every depth at once).  The
  // perturbed copies of the chunk run side by side as one batch per group:
  // group 0 is T (whose continuum and chi_ref tangents are taken too), group
  // 1 the other parameters; copy 2j is parameter j of the group shifted up,
  // 2j + 1 shifted down.  Each column's state does not depend on its
  // neighbours, so the batch gives the same values as one launch per copy.
  const int64_t natm = (int64_t) ncol * 9 * Nmax;
  const int64_t nst = npts * nfield;
  double *dstate = (double *) c->rf_dstate.p;
  int32_t *rep_ndep = (int32_t *) c->rf_ndep.p;
  {
    dim3 blk(256), grd((unsigned) ((ncol + 255) / 256), (unsigned) (2 * pl.npar));
    hipLaunchKernelGGL(rhg::k_rf_rep_ndep, grd, blk, 0, c->stream, d_ndep, ncol, rep_ndep);
    HIPCHECK(hipGetLastError());
  }
  for (int grp = 0; grp < 2; grp++) {
    int gp[RHG_RF_NPARAM], ng = 0;
    for (int ip = 0; ip < pl.npar; ip++)
      if ((pl.p[ip] == RHG_RF_T) == (grp == 0)) gp[ng++] = pl.p[ip];
    if (ng == 0) continue;
    rhg::RfShiftList shift{};
    for (int j = 0; j < ng; j++) {
      shift.row[2 * j] = shift.row[2 * j + 1] = kRfParam[gp[j]].row;
      shift.h[2 * j] = kRfParam[gp[j]].h;
      shift.h[2 * j + 1] = -kRfParam[gp[j]].h;
    }
    const int W = 2 * ng * ncol;
    double *atm_b = (double *) c->rf_atm[grp].p, *st_b = (double *) c->rf_state[grp].p;
    {
      dim3 blk(256), grd((unsigned) ((natm + 255) / 256), (unsigned) (2 * ng));
      hipLaunchKernelGGL(rhg::k_rf_perturb, grd, blk, 0, c->stream, d_atm, atm_b, natm, Nmax, shift);
      HIPCHECK(hipGetLastError());
    }
    const bool gT = grp == 0;
    rc = launch_base<double>(c, atm_b, rep_ndep, W, Nmax, mu, scale, st_b, nfield,
                             (int32_t *) c->rf_moving.p, gT ? (double *) c->rf_chiref[0].p : nullptr,
                             gT ? (double *) c->rf_cont[0].p : nullptr,
                             gT ? (double *) c->rf_cont[1].p : nullptr);
    if (rc) return rc;
    dim3 blk(256), grd((unsigned) ((nst + 255) / 256));
    for (int j = 0; j < ng; j++) {
      hipLaunchKernelGGL(rhg::k_rf_diff<double>, grd, blk, 0, c->stream, (const double *) st_b,
                         (int64_t) W, (int64_t) 2 * j * ncol, (int64_t) (2 * j + 1) * ncol, ncol,
                         (int64_t) nfield * Nmax, 0.5 / kRfParam[gp[j]].h, dstate + gp[j] * nst);
      HIPCHECK(hipGetLastError());
    }
    if (gT) {
      const double inv2h = 0.5 / kRfParam[RHG_RF_T].h;
      const int64_t ncc = npts * nout;
      dim3 grdc((unsigned) ((ncc + 255) / 256));
      for (int q = 0; q < 2; q++) {
        hipLaunchKernelGGL(rhg::k_rf_diff<R>, grdc, blk, 0, c->stream, (const double *) c->rf_cont[q].p,
                           (int64_t) W, (int64_t) 0, (int64_t) ncol, ncol, (int64_t) nout * Nmax,
                           inv2h, (R *) c->rf_dcont[q].p);
        HIPCHECK(hipGetLastError());
      }
      // chi_ref is [column][depth]: the two copies are the halves
      dim3 grd2((unsigned) ((npts + 255) / 256));
      hipLaunchKernelGGL(rhg::k_rf_diff<double>, grd2, blk, 0, c->stream,
                         (const double *) c->rf_chiref[0].p, (int64_t) 0, (int64_t) 0, npts,
                         (int) npts, (int64_t) 1, inv2h, (double *) c->rf_dchiref.p);
      HIPCHECK(hipGetLastError());
    }
  }
  HIPCHECK(hipEventRecord(c->ev[2], c->stream));

  // 2-4. tape, reverse sweep, chain (polarized wavelengths); the others by
  // the scalar solver with zero response functions
  const int ncb = (ncol + 63) / 64;
  dim3 blk(64);
  if (npol > 0) {
    dim3 grd((unsigned) rhg::xcd_grid(ncb, npol));
    double *car = (double *) c->rf_car.p;
    R *aP = (R *) c->rf_aP.p, *aZ = (R *) c->rf_aZ.p, *aB = (R *) c->rf_aB.p;
    hipLaunchKernelGGL(rhg::k_rf_forward<R>, grd, blk, 0, c->stream, c->d_model, d_ndep, moving, ncol,
                       Nmax, state, c->d_out_ls, c->d_io_pol, npol, nout, cchi, ceta, mu, d_out,
                       car);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(c->ev[3], c->stream));
    hipLaunchKernelGGL(rhg::k_rf_reverse<R>, grd, blk, 0, c->stream, c->d_model, d_ndep, moving, ncol,
                       Nmax, state, c->d_out_ls, c->d_io_pol, npol, cchi, ceta, mu, car, aP, aZ,
                       aB);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(c->ev[4], c->stream));
    if (pl.npar <= 3)
      hipLaunchKernelGGL((rhg::k_rf_chain<R, 3>), grd, blk, 0, c->stream, c->d_model, d_ndep, moving,
                         ncol, Nmax, state, dstate, chiref, (const double *) c->rf_dchiref.p, cchi,
                         ceta, (const R *) c->rf_dcont[0].p, (const R *) c->rf_dcont[1].p,
                         c->d_out_ls, c->d_io_pol, npol, nout, mu, scale, pl, aP, aZ, aB, d_rf, ldr);
    else
      hipLaunchKernelGGL((rhg::k_rf_chain<R, 6>), grd, blk, 0, c->stream, c->d_model, d_ndep, moving,
                         ncol, Nmax, state, dstate, chiref, (const double *) c->rf_dchiref.p, cchi,
                         ceta, (const R *) c->rf_dcont[0].p, (const R *) c->rf_dcont[1].p,
                         c->d_out_ls, c->d_io_pol, npol, nout, mu, scale, pl, aP, aZ, aB, d_rf, ldr);
    HIPCHECK(hipGetLastError());
  } else {
    HIPCHECK(hipEventRecord(c->ev[3], c->stream));
    HIPCHECK(hipEventRecord(c->ev[4], c->stream));
  }
  if (c->n_scal > 0) {
    dim3 grd((unsigned) rhg::xcd_grid(ncb, c->n_scal));
    hipLaunchKernelGGL(rhg::k_formal_scalar<R>, grd, blk, 0, c->stream, c->d_model, d_ndep, moving,
                       ncol, Nmax, state, c->d_out_ls, c->d_io_scal, c->n_scal, nout,
                       (const R *) cchi, (const R *) ceta, mu, d_out, (const int32_t *) nullptr, (const int32_t *) nullptr, (const int32_t *) nullptr);
    HIPCHECK(hipGetLastError());
  }
  HIPCHECK(hipEventRecord(c->ev[5], c->stream));
  HIPCHECK(hipEventSynchronize(c->ev[5]));
  float ms;
  for (int st = 0; st < 5; st++) {
    HIPCHECK(hipEventElapsedTime(&ms, c->ev[st], c->ev[st + 1]));
    c->rf_ms[st] += ms;
  }
  return 0;
}

// The response functions of a batch, in column chunks whose work buffers
// (column state, continuum, tangents, the forward sweep's carried state and
// the point adjoints: about npol * Nmax * 300 bytes per column) fit in half
// of the device memory available; the outputs are the whole batch's.
template <typename R>
int rf_run(rhg_context_t *c, const rhg_columns_t *cols, int32_t device_ptrs,
           const rhg::RfParamList &pl, double *stokes_out, void *rf_out) {
  const int ncol_all = cols->ncol, Nmax = cols->ndep_max, nout = c->nout;
  const int nfield = rhg::state_nfield(c->ncont, c->nline);
  const size_t npl = (size_t) (c->n_pol > 0 ? c->n_pol : 1);
  const size_t rsz = sizeof(R);
  // the perturbed copies of the state tangents: 2 for T, 2 per other parameter
  int nT = 0;
  for (int i = 0; i < pl.npar; i++) nT += pl.p[i] == RHG_RF_T;
  const size_t ncopy = 2 * (size_t) pl.npar, nother = (size_t) (pl.npar - nT);
  const size_t per_col =
      (size_t) Nmax * ((size_t) nfield * 8 + 2 * nout * rsz + 8 + ncopy * 9 * 8 +
                       ncopy * (size_t) nfield * 8 + 4 * (size_t) nout * 8 + 2 * 8 +
                       RHG_RF_NPARAM * (size_t) nfield * 8 + 2 * nout * rsz + 8 +
                       npl * rhg::RF_NCARRY * 8 + npl * 4 * 8 * rsz + npl * 4 * rsz) +
      npl * 8 * rsz + (1 + 2 * ncopy) * sizeof(int32_t);
  DevBuf *work[] = {&c->state, &c->moving, &c->cchi, &c->ceta, &c->chiref, &c->rf_atm[0],
                    &c->rf_atm[1], &c->rf_state[0], &c->rf_state[1], &c->rf_cont[0],
                    &c->rf_cont[1], &c->rf_cont[2], &c->rf_cont[3], &c->rf_chiref[0],
                    &c->rf_chiref[1], &c->rf_moving, &c->rf_ndep, &c->rf_dstate, &c->rf_dcont[0],
                    &c->rf_dcont[1], &c->rf_dchiref, &c->rf_car, &c->rf_aP, &c->rf_aZ, &c->rf_aB};
  size_t held = 0;
  for (DevBuf *b : work) held += b->bytes;
  size_t freeb = 0, totb = 0;
  HIPCHECK(hipMemGetInfo(&freeb, &totb));
  const size_t out_b = (size_t) ncol_all * 4 * nout * sizeof(double);
  const size_t rf_b = (size_t) pl.npar * Nmax * 4 * nout * ncol_all * rsz;
  const size_t atm_all_b = (size_t) ncol_all * 9 * Nmax * sizeof(double);
  size_t avail = freeb + held;
  if (!device_ptrs) {
    const size_t host_side = out_b + rf_b + atm_all_b + ncol_all * sizeof(int32_t);
    avail = avail > host_side ? avail - host_side : 0;
  }
  int64_t chunk = (int64_t) (avail / 2 / per_col);
  if (chunk < 1) return fail(RHG_ENOMEM, "device memory too small for one response-function column");
  if (chunk > ncol_all) chunk = ncol_all;
  if (chunk > 64) chunk = chunk / 64 * 64;
  // PYRH_AMD_RF_CHUNK caps the chunk (tests of the chunked path)
  if (const char *cap = getenv("PYRH_AMD_RF_CHUNK")) {
    const long v = strtol(cap, nullptr, 10);
    if (v >= 1 && v < chunk) chunk = v;
  }
  const int64_t npts = chunk * Nmax;
  const size_t atm_b = (size_t) chunk * 9 * Nmax * sizeof(double);
  const size_t st_b = (size_t) npts * nfield * sizeof(double);
  const size_t cc_b = (size_t) npts * nout * rsz;
  const size_t cc64_b = (size_t) npts * nout * sizeof(double);
  const size_t car_b = npl * Nmax * rhg::RF_NCARRY * chunk * sizeof(double);
  const size_t aP_b = npl * 4 * Nmax * 8 * chunk * rsz;
  const size_t aZ_b = npl * 4 * Nmax * chunk * rsz;
  const size_t aB_b = npl * 4 * 2 * chunk * rsz;
  if (c->state.ensure(st_b) || c->moving.ensure(chunk * sizeof(int32_t)) || c->cchi.ensure(cc_b) ||
      c->ceta.ensure(cc_b) || c->chiref.ensure((size_t) npts * sizeof(double)) ||
      (!device_ptrs && (c->atmos.ensure(atm_all_b) || c->ndep.ensure(ncol_all * sizeof(int32_t)) ||
                        c->stokes.ensure(out_b) || c->rf_out.ensure(rf_b))) ||
      c->rf_atm[0].ensure(2 * nT * atm_b) || c->rf_atm[1].ensure(2 * nother * atm_b) ||
      c->rf_state[0].ensure(2 * nT * st_b) || c->rf_state[1].ensure(2 * nother * st_b) ||
      c->rf_cont[0].ensure(2 * nT * cc64_b) || c->rf_cont[1].ensure(2 * nT * cc64_b) ||
      c->rf_chiref[0].ensure(2 * nT * (size_t) npts * sizeof(double)) ||
      c->rf_moving.ensure(ncopy * chunk * sizeof(int32_t)) ||
      c->rf_ndep.ensure(ncopy * chunk * sizeof(int32_t)) ||
      c->rf_dstate.ensure(st_b * RHG_RF_NPARAM) || c->rf_dcont[0].ensure(cc_b) ||
      c->rf_dcont[1].ensure(cc_b) || c->rf_dchiref.ensure((size_t) npts * sizeof(double)) ||
      c->rf_car.ensure(car_b) || c->rf_aP.ensure(aP_b) ||
      c->rf_aZ.ensure(aZ_b) || c->rf_aB.ensure(aB_b))
    return fail(RHG_ENOMEM, "device allocation for the response functions failed");
  const double *d_atm = (const double *) cols->atmos;
  const int32_t *d_ndep = cols->ndep;
  double *d_out = stokes_out;
  R *d_rf = (R *) rf_out;
  if (!device_ptrs) {
    HIPCHECK(hipMemcpyAsync(c->atmos.p, cols->atmos, atm_all_b, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->ndep.p, cols->ndep, ncol_all * sizeof(int32_t), hipMemcpyHostToDevice,
                            c->stream));
    d_atm = (const double *) c->atmos.p;
    d_ndep = (const int32_t *) c->ndep.p;
    d_out = (double *) c->stokes.p;
    d_rf = (R *) c->rf_out.p;
  }
  // the wavelengths of the scalar solvers keep zero response functions
  HIPCHECK(hipMemsetAsync(d_rf, 0, rf_b, c->stream));
  for (float &x : c->rf_ms) x = 0.0f;
  for (int64_t c0 = 0; c0 < ncol_all; c0 += chunk) {
    const int n = (int) std::min<int64_t>(chunk, ncol_all - c0);
    if (int rc = rf_chunk<R>(c, d_atm + c0 * 9 * Nmax, d_ndep + c0, n, Nmax, cols->mu,
                             cols->atm_scale, pl, d_out + c0 * 4 * nout, d_rf + c0, ncol_all))
      return rc;
  }
  if (!device_ptrs) {
    HIPCHECK(hipMemcpyAsync(stokes_out, d_out, out_b, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipMemcpyAsync(rf_out, d_rf, rf_b, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  // rhg_last_timings' four intervals: inputs, tangents, forward, reverse + chain
  for (int st = 0; st < 3; st++) c->last_ms[st] = c->rf_ms[st];
  c->last_ms[3] = c->rf_ms[3] + c->rf_ms[4];
  return 0;
}

}  // namespace

extern "C" {

int rhg_rf_lte(rhg_context_t *c, const rhg_columns_t *cols, int32_t precision, int32_t device_ptrs,
               double *stokes_out, void *rf_out) {
  const int32_t all[RHG_RF_NPARAM] = {RHG_RF_T, RHG_RF_VLOS, RHG_RF_B, RHG_RF_GAMMA, RHG_RF_CHI,
                                      RHG_RF_VMIC};
  return rhg_rf_lte_params(c, cols, precision, device_ptrs, RHG_RF_NPARAM, all, stokes_out, rf_out);
}

int rhg_rf_lte_params(rhg_context_t *c, const rhg_columns_t *cols, int32_t precision,
                      int32_t device_ptrs, int32_t nparam, const int32_t *params,
                      double *stokes_out, void *rf_out) {
  if (!c || !cols || !stokes_out || !rf_out || !params) return fail(RHG_EINVAL, "null argument");
  if (nparam < 1 || nparam > RHG_RF_NPARAM) return fail(RHG_EINVAL, "bad number of parameters");
  if (c->dm.magneto_optical)
    return fail(RHG_EUNSUPPORTED,
                "response functions with the magneto-optical terms are not implemented");
  rhg::RfParamList pl{};
  pl.npar = nparam;
  for (int i = 0; i < nparam; i++) {
    if (params[i] < 0 || params[i] >= RHG_RF_NPARAM) return fail(RHG_EINVAL, "bad parameter id");
    for (int j = 0; j < i; j++)
      if (params[j] == params[i]) return fail(RHG_EINVAL, "repeated parameter id");
    pl.p[i] = params[i];
  }
  if (c->nout < 1) return fail(RHG_EINVAL, "the model has no output wavelength");
  if (precision != RHG_PREC_FP64 && precision != RHG_PREC_FP32)
    return fail(RHG_EINVAL, "precision must be RHG_PREC_FP64 or RHG_PREC_FP32");
  const int ncol = cols->ncol, Nmax = cols->ndep_max;
  if (ncol < 1 || ncol > RHG_RF_MAX_COLUMNS || Nmax < 4 || Nmax > RHG_MAX_DEPTH)
    return fail(RHG_EINVAL, "bad batch shape (at most RHG_RF_MAX_COLUMNS columns per call)");
  if (!(cols->mu > 0.0 && cols->mu <= 1.0)) return fail(RHG_EINVAL, "mu must be in (0, 1]");
  if (cols->atm_scale < 0 || cols->atm_scale > 2) return fail(RHG_EINVAL, "bad atm_scale");
  std::vector<int32_t> hndep(ncol);
  if (device_ptrs) {
    HIPCHECK(hipMemcpy(hndep.data(), cols->ndep, ncol * sizeof(int32_t), hipMemcpyDeviceToHost));
  } else {
    memcpy(hndep.data(), cols->ndep, ncol * sizeof(int32_t));
  }
  for (int i = 0; i < ncol; i++)
    if (hndep[i] < 4 || hndep[i] > Nmax) return fail(RHG_EINVAL, "ndep out of range");
  HIPCHECK(hipSetDevice(c->device));
  std::vector<int32_t> perm, sndep;
  if (int rc = sort_order(c, cols, device_ptrs, perm, sndep)) return rc;
  if (perm.empty()) {
    const int rc = precision == RHG_PREC_FP32
                       ? rf_run<float>(c, cols, device_ptrs, pl, stokes_out, rf_out)
                       : rf_run<double>(c, cols, device_ptrs, pl, stokes_out, rf_out);
    if (rc) return rc;
    g_err.clear();
    return 0;
  }
  // uneven depth counts: run sorted by depth count, scatter both outputs back
  rhg_columns_t cols2;
  if (int rc = gather_sorted(c, cols, device_ptrs, perm, sndep, cols2)) return rc;
  const size_t esz = precision == RHG_PREC_FP32 ? sizeof(float) : sizeof(double);
  const int64_t nouter = (int64_t) pl.npar * Nmax * 4 * c->nout;
  const size_t rf_b = (size_t) nouter * ncol * esz;
  if (c->srt_out.ensure((size_t) ncol * 4 * c->nout * sizeof(double)) || c->srt_rf.ensure(rf_b) ||
      (!device_ptrs && c->rf_out.ensure(rf_b)))
    return fail(RHG_ENOMEM, "device allocation for the sorted batch failed");
  int rc = precision == RHG_PREC_FP32
               ? rf_run<float>(c, &cols2, 1, pl, (double *) c->srt_out.p, c->srt_rf.p)
               : rf_run<double>(c, &cols2, 1, pl, (double *) c->srt_out.p, c->srt_rf.p);
  if (rc) return rc;
  if ((rc = scatter_stokes(c, ncol, device_ptrs, (const double *) c->srt_out.p, stokes_out))) return rc;
  void *rdst = device_ptrs ? rf_out : c->rf_out.p;
  const dim3 g((unsigned) ((nouter * ncol + 255) / 256)), b(256);
  if (precision == RHG_PREC_FP32)
    hipLaunchKernelGGL(rhg::k_scatter_last<float>, g, b, 0, c->stream, (const float *) c->srt_rf.p,
                       (float *) rdst, (const int32_t *) c->srt_perm.p, ncol, nouter);
  else
    hipLaunchKernelGGL(rhg::k_scatter_last<double>, g, b, 0, c->stream, (const double *) c->srt_rf.p,
                       (double *) rdst, (const int32_t *) c->srt_perm.p, ncol, nouter);
  HIPCHECK(hipGetLastError());
  if (!device_ptrs) HIPCHECK(hipMemcpyAsync(rf_out, rdst, rf_b, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  g_err.clear();
  return 0;
}

int rhg_rf_loggf(rhg_context_t *c, const rhg_columns_t *cols, int32_t npar,
                 const int32_t *par_line, int32_t device_ptrs, double *stokes_out,
                 double *rf_out) {
  if (!c || !cols || !stokes_out || !rf_out || (npar > 0 && !par_line))
    return fail(RHG_EINVAL, "null argument");
  if (c->nout < 1) return fail(RHG_EINVAL, "the model has no output wavelength");
  if (npar < 1 || npar > RHG_RF_MAX_ATOMIC)
    return fail(RHG_EINVAL, "npar must be in [1, RHG_RF_MAX_ATOMIC]");
  for (int p = 0; p < npar; p++)
    if (par_line[p] >= c->nline) return fail(RHG_EINVAL, "par_line entry beyond the model's lines");
  const int ncol = cols->ncol, Nmax = cols->ndep_max, nout = c->nout;
  if (ncol < 1 || ncol > RHG_RF_MAX_COLUMNS)
    return fail(RHG_EINVAL, "bad batch shape (at most RHG_RF_MAX_COLUMNS columns per call)");
  // the synthesis leaves the column state, the FP64 continuum and the
  // depth counts of this batch in the context
  int rc = synth(c, cols, RHG_PREC_FP64, device_ptrs, false, stokes_out, nullptr, nullptr);
  if (rc) return rc;
  const size_t rf_b = (size_t) ncol * nout * npar * sizeof(double);
  const int nscal = c->n_scal > 0 ? c->n_scal : 1;
  if (c->at_iin.ensure((size_t) nscal * Nmax * ncol * sizeof(double)) ||
      c->at_par.ensure((size_t) npar * sizeof(int32_t)) || (!device_ptrs && c->at_out.ensure(rf_b)))
    return fail(RHG_ENOMEM, "device allocation for the log gf response functions failed");
  double *d_rf = device_ptrs ? rf_out : (double *) c->at_out.p;
  HIPCHECK(hipMemcpyAsync(c->at_par.p, par_line, npar * sizeof(int32_t), hipMemcpyHostToDevice,
                          c->stream));
  HIPCHECK(hipMemsetAsync(d_rf, 0, rf_b, c->stream));
  if (c->n_scal > 0) {
    dim3 blk(64), grd((unsigned) rhg::xcd_grid((ncol + 63) / 64, c->n_scal));
    hipLaunchKernelGGL(rhg::k_rf_loggf, grd, blk, 0, c->stream, c->d_model, c->last_ndep,
                       (const int32_t *) c->moving.p, ncol, Nmax, (const double *) c->state.p,
                       c->d_out_ls, c->d_io_scal, c->n_scal, nout, (const double *) c->cchi.p,
                       (const double *) c->ceta.p, cols->mu, npar, (const int32_t *) c->at_par.p,
                       (double *) c->at_iin.p, d_rf);
    HIPCHECK(hipGetLastError());
  }
  if (!device_ptrs)
    HIPCHECK(hipMemcpyAsync(rf_out, d_rf, rf_b, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  g_err.clear();
  return 0;
}

}  // extern "C"

namespace {

// Host-side checks of an active atom against the model of the context and
// what the NLTE kernels implement; fills the derived tables.
struct NlteHost {
  // output-wavelength positions with an active set: those with a line
  // (Bezier rays) first, then the continuum-only ones (Feautrier)
  std::vector<int32_t> ln_tr, ln_phi, io_list;
  int nbez = 0, nentry = 0, nphi = 0, maxset_bez = 0, maxset_fe = 0;
};

int nlte_check(const rhg_context_t *c, const rhg_nlte_t *a, NlteHost &h) {
  const int ns = c->nspect, nl = a->nlevel;
  if (nl < 2 || nl > rhg::NLTE_MAXLEV) return fail(RHG_EUNSUPPORTED, "active atom: 2..16 levels");
  if (a->atom_index < 0 || a->atom_index >= c->natom) return fail(RHG_EINVAL, "atom_index out of range");
  if (c->h_atom_level_ptr[a->atom_index + 1] - c->h_atom_level_ptr[a->atom_index] != nl)
    return fail(RHG_EINVAL, "active atom: level count differs from the model's atom");
  if (a->he_atom >= c->natom) return fail(RHG_EINVAL, "he_atom out of range");
  for (int32_t ia : c->h_nucl_atom)
    if (ia == a->atom_index || ia == a->he_atom)
      return fail(RHG_EUNSUPPORTED, "active atom or He bound in molecules");
  if (a->nrays < 1 || a->nrays > 64) return fail(RHG_EINVAL, "nrays out of range");
  if (a->ng_order < 0 || a->ng_order > rhg::NLTE_MAXNG) return fail(RHG_EUNSUPPORTED, "Ng order > 4");
  if (a->ng_order > 0 && a->ng_period < 1) return fail(RHG_EINVAL, "Ng period < 1");
  if (a->isum < -1 || a->isum >= nl) return fail(RHG_EINVAL, "isum out of range");
  if (a->nmax_iter < 0 || a->nmax_scatter < 0) return fail(RHG_EINVAL, "negative iteration count");
  if (a->ntrans < 1 || a->nline < 0) return fail(RHG_EINVAL, "no transitions");
  h.ln_tr.assign(a->nline > 0 ? a->nline : 1, -1);
  h.ln_phi.assign(a->nline > 0 ? a->nline : 1, 0);
  if (a->tr_ptr[0] != 0) return fail(RHG_EINVAL, "tr_ptr[0] != 0");
  for (int t = 0; t < a->ntrans; t++) {
    const int i = a->tr_i[t], j = a->tr_j[t];
    if (i < 0 || j <= i || j >= nl) return fail(RHG_EINVAL, "transition levels out of range");
    if (a->tr_nblue[t] < 0 || a->tr_nlam[t] < 1 || a->tr_nblue[t] + a->tr_nlam[t] > ns)
      return fail(RHG_EINVAL, "transition span outside the grid");
    if (a->tr_ptr[t + 1] - a->tr_ptr[t] != a->tr_nlam[t]) return fail(RHG_EINVAL, "tr_ptr inconsistent");
    if (a->tr_type[t] == 0) {
      const int l = a->tr_line[t];
      if (l < 0 || l >= a->nline || h.ln_tr[l] >= 0) return fail(RHG_EINVAL, "bad tr_line");
      h.ln_tr[l] = t;
      h.ln_phi[l] = h.nphi;
      h.nphi += a->tr_nlam[t];
    } else if (a->tr_type[t] != 1) {
      return fail(RHG_EINVAL, "transition type must be 0 (line) or 1 (continuum)");
    }
  }
  for (int l = 0; l < a->nline; l++)
    if (h.ln_tr[l] < 0) return fail(RHG_EINVAL, "line without a transition");
  h.nentry = a->tr_ptr[a->ntrans];
  if (a->as_ptr[0] != 0) return fail(RHG_EINVAL, "as_ptr[0] != 0");
  for (int ls = 0; ls < ns; ls++) {
    const int n = a->as_ptr[ls + 1] - a->as_ptr[ls];
    if (n < 0 || n > rhg::NLTE_MAXSET) return fail(RHG_EUNSUPPORTED, "more than 8 transitions active at one wavelength");
    for (int e = a->as_ptr[ls]; e < a->as_ptr[ls + 1]; e++) {
      const int t = a->as_tr[e];
      if (t < 0 || t >= a->ntrans || ls < a->tr_nblue[t] || ls >= a->tr_nblue[t] + a->tr_nlam[t])
        return fail(RHG_EINVAL, "active set entry outside its transition's span");
    }
    if (n > 0 && ls == c->dm.ref_index)
      return fail(RHG_EUNSUPPORTED, "active transition at the reference wavelength");
    if (n > 0 && a->as_bezier[ls]) {
      h.io_list.push_back(ls < c->dm.ref_index ? ls : ls - 1);
      h.maxset_bez = std::max(h.maxset_bez, n);
    }
  }
  h.nbez = (int) h.io_list.size();
  // every other output wavelength, active or not: J is iterated there too,
  // as the reference solves every wavelength in each pass
  for (int ls = 0; ls < ns; ls++)
    if (ls != c->dm.ref_index && !(a->as_ptr[ls + 1] > a->as_ptr[ls] && a->as_bezier[ls])) {
      h.io_list.push_back(ls < c->dm.ref_index ? ls : ls - 1);
      h.maxset_fe = std::max(h.maxset_fe, a->as_ptr[ls + 1] - a->as_ptr[ls]);
    }
  for (int t = 0; t < a->ntrans; t++)
    for (int ls = a->tr_nblue[t]; ls < a->tr_nblue[t] + a->tr_nlam[t]; ls++) {
      bool found = false;
      for (int e = a->as_ptr[ls]; e < a->as_ptr[ls + 1]; e++) found |= a->as_tr[e] == t;
      if (!found) return fail(RHG_EINVAL, "transition missing from an active set in its span");
    }
  for (int r = 0; r < a->ncoll; r++) {
    const int n = a->coll_ptr[r + 1] - a->coll_ptr[r];
    if (n < 2) return fail(RHG_EINVAL, "collision table with fewer than 2 temperatures");
    for (int q = 1; q < n; q++)
      if (!(a->coll_T[a->coll_ptr[r] + q] > a->coll_T[a->coll_ptr[r] + q - 1]))
        return fail(RHG_EINVAL, "collision temperatures must ascend");
    if (a->coll_i[r] < 0 || a->coll_i[r] >= nl || a->coll_j[r] < 0 || a->coll_j[r] >= nl ||
        a->coll_type[r] < 0 || a->coll_type[r] > 6)
      return fail(RHG_EINVAL, "bad collision entry");
  }
  if (h.nbez == 0 && a->as_ptr[ns] == 0) return fail(RHG_EINVAL, "no active wavelength");
  return 0;
}

// Copy the atom into one device buffer and point a DevNlte at it.
int nlte_upload(rhg_context_t *c, const rhg_nlte_t *a, const NlteHost &h) {
  std::vector<char> blob;
  std::vector<std::pair<size_t, const void **>> fix;
  rhg::DevNlte d{};
  auto add = [&](const void *src, size_t n, size_t esz, const void **dst) {
    size_t off = (blob.size() + 255) & ~(size_t) 255;
    blob.resize(off + (n > 0 ? n : 1) * esz, 0);
    if (n > 0 && src) memcpy(blob.data() + off, src, n * esz);
    fix.push_back({off, dst});
  };
#define ADD(field, n) add(a->field, (size_t) (n), sizeof(*a->field), (const void **) &d.field)
  const int nl = a->nlevel, nt = a->ntrans, nln = a->nline, ns = c->nspect;
  const int ncw = a->ncoll > 0 ? a->coll_ptr[a->ncoll] : 0;
  ADD(E, nl); ADD(g, nl); ADD(stage, nl);
  ADD(tr_type, nt); ADD(tr_i, nt); ADD(tr_j, nt); ADD(tr_nblue, nt); ADD(tr_nlam, nt);
  ADD(tr_ptr, nt + 1); ADD(tr_line, nt); ADD(tr_wlam, h.nentry); ADD(tr_alpha, h.nentry);
  ADD(ln_lambda0, nln); ADD(ln_Aji, nln); ADD(ln_Bji, nln); ADD(ln_Bij, nln); ADD(ln_Grad, nln);
  ADD(ln_vdw_cross, nln); ADD(ln_vdw_alpha, nln); ADD(ln_vdw_cross_he, nln);
  ADD(ln_stark_c, nln); ADD(ln_stark_C, nln); ADD(ln_stark_Cm, nln);
  ADD(ln_vdw, nln); ADD(ln_stark, nln);
  ADD(as_ptr, ns + 1); ADD(as_tr, a->as_ptr[ns]); ADD(as_bezier, ns);
  ADD(coll_type, a->ncoll); ADD(coll_i, a->ncoll); ADD(coll_j, a->ncoll); ADD(coll_ptr, a->ncoll + 1);
  ADD(coll_T, ncw); ADD(coll_c, ncw); ADD(coll_M, ncw);
  ADD(muz, a->nrays); ADD(wmu, a->nrays); ADD(ray_H, ns); ADD(ray_He, ns); ADD(ray_H2, ns);
#undef ADD
  add(h.ln_tr.data(), h.ln_tr.size(), sizeof(int32_t), (const void **) &d.ln_tr);
  add(h.ln_phi.data(), h.ln_phi.size(), sizeof(int32_t), (const void **) &d.ln_phi);
  if (c->nl_atom.ensure(blob.size()) || c->nl_dev.ensure(sizeof(rhg::DevNlte)))
    return fail(RHG_ENOMEM, "device allocation for the active atom failed");
  for (auto &f : fix) *f.second = (const char *) c->nl_atom.p + f.first;
  d.nl = nl;
  d.atom_index = a->atom_index;
  d.he_atom = a->he_atom;
  d.ntrans = nt;
  d.nline = nln;
  d.ncoll = a->ncoll;
  d.nrays = a->nrays;
  d.nentry = h.nentry;
  d.nphi = h.nphi;
  d.weight = a->weight;
  d.sigma_thomson = a->sigma_thomson;
  HIPCHECK(hipMemcpyAsync(c->nl_atom.p, blob.data(), blob.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->nl_dev.p, &d, sizeof(d), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

struct NlteTimer {
  hipEvent_t e[2]{};
  ~NlteTimer() {
    for (auto &x : e)
      if (x) (void) hipEventDestroy(x);
  }
};

}  // namespace

extern "C" {

int rhg_nlte_solve(rhg_context_t *c, const rhg_nlte_t *a, const rhg_columns_t *cols,
                   int32_t device_ptrs, double *pops_out, int32_t *niter_out,
                   double *dpops_out, double *stokes_out, double *nstar_out) {
  if (!c || !a || !cols || !pops_out || !niter_out || !dpops_out)
    return fail(RHG_EINVAL, "null argument");
  if (c->nout < 1) return fail(RHG_EINVAL, "the model has no output wavelength");
  const int ncol_all = cols->ncol, Nmax = cols->ndep_max, nout = c->nout, nl = a->nlevel;
  if (ncol_all < 1 || Nmax < 4 || Nmax > RHG_MAX_DEPTH) return fail(RHG_EINVAL, "bad batch shape");
  if (!(cols->mu > 0.0 && cols->mu <= 1.0)) return fail(RHG_EINVAL, "mu must be in (0, 1]");
  if (cols->atm_scale < 0 || cols->atm_scale > 2) return fail(RHG_EINVAL, "bad atm_scale");
  NlteHost h;
  if (int rc = nlte_check(c, a, h)) return rc;
  std::vector<int32_t> hndep(ncol_all);
  HIPCHECK(hipSetDevice(c->device));
  if (device_ptrs) {
    HIPCHECK(hipMemcpy(hndep.data(), cols->ndep, ncol_all * sizeof(int32_t), hipMemcpyDeviceToHost));
  } else {
    memcpy(hndep.data(), cols->ndep, ncol_all * sizeof(int32_t));
  }
  for (int i = 0; i < ncol_all; i++)
    if (hndep[i] < 4 || hndep[i] > Nmax) return fail(RHG_EINVAL, "ndep out of range");
  if (int rc = nlte_upload(c, a, h)) return rc;
  const rhg::DevNlte *d_at = (const rhg::DevNlte *) c->nl_dev.p;
  const int nlist = (int) h.io_list.size(), nrays = a->nrays, NS = a->ng_order + 2;
  const int nfield = rhg::state_nfield(c->ncont, c->nline);

  // device bytes per column; columns are solved in chunks that fit in 3/4
  // of the free device memory
  const size_t per_col =
      (size_t) Nmax * sizeof(double) *
          ((size_t) nfield + 5 * (size_t) nout + 2 * (size_t) nl + 3 + 2 * (size_t) a->nline +
           2 * (size_t) nl * nl + (size_t) h.nphi * nrays * 2 + 2 * (size_t) h.nentry +
           std::max((size_t) rhg::NLTE_NSLOT * (nlist - h.nbez),
                    stokes_out ? (size_t) rhg::NLTE_NSLOT_EM * nlist : (size_t) 0) +
           (size_t) NS * nl +
           (size_t) nout) +
      (size_t) nlist * sizeof(double) + 9 * (size_t) Nmax * sizeof(double) + 64;
  size_t freeb = 0, totb = 0;
  HIPCHECK(hipMemGetInfo(&freeb, &totb));
  const size_t budget = freeb / 4 * 3;
  int chunk = (int) std::min<size_t>((size_t) ncol_all, budget / per_col);
  if (chunk > 64) chunk = chunk / 64 * 64;
  if (chunk < 1) return fail(RHG_ENOMEM, "device memory too small for one NLTE column");

  NlteTimer tm_all, tm_k;
  for (auto &x : tm_all.e) HIPCHECK(hipEventCreate(&x));
  for (auto &x : tm_k.e) HIPCHECK(hipEventCreate(&x));
  for (double &x : c->nlte_ms) x = 0.0;
  HIPCHECK(hipEventRecord(tm_all.e[0], c->stream));
  float ms = 0.0f;
  // time one launch sequence on the stream into nlte_ms[slot]
  auto timed = [&](int slot, auto &&launch) -> int {
    HIPCHECK(hipEventRecord(tm_k.e[0], c->stream));
    if (int rc = launch()) return rc;
    HIPCHECK(hipEventRecord(tm_k.e[1], c->stream));
    HIPCHECK(hipEventSynchronize(tm_k.e[1]));
    HIPCHECK(hipEventElapsedTime(&ms, tm_k.e[0], tm_k.e[1]));
    c->nlte_ms[slot] += ms;
    return 0;
  };

  if (c->nl_iolist.ensure((size_t) nlist * sizeof(int32_t)))
    return fail(RHG_ENOMEM, "device allocation (NLTE wavelength list)");
  HIPCHECK(hipMemcpyAsync(c->nl_iolist.p, h.io_list.data(), nlist * sizeof(int32_t),
                          hipMemcpyHostToDevice, c->stream));
  const int32_t *d_iolist = (const int32_t *) c->nl_iolist.p;

  for (int c0 = 0; c0 < ncol_all; c0 += chunk) {
    const int ncol = std::min(chunk, ncol_all - c0);
    const size_t pts = (size_t) ncol * Nmax * sizeof(double);
    const int ncb_max = (ncol + 63) / 64;
    if (c->state.ensure(pts * nfield) || c->cchi.ensure(pts * nout) || c->ceta.ensure(pts * nout) ||
        c->moving.ensure(ncol * sizeof(int32_t)) || c->nl_nstar.ensure(pts * nl) ||
        c->nl_pops.ensure(pts * nl) || c->nl_ntot.ensure(pts) || c->nl_vbroad.ensure(pts) ||
        c->nl_nHe.ensure(pts) || c->nl_adamp.ensure(pts * std::max(a->nline, 1)) ||
        c->nl_wphi.ensure(pts * std::max(a->nline, 1)) || c->nl_Cmat.ensure(pts * nl * nl) ||
        c->nl_Gam.ensure(pts * nl * nl) ||
        c->nl_phi.ensure(pts * std::max(h.nphi, 1) * nrays * 2) || c->nl_csca.ensure(pts * nout) ||
        c->nl_J.ensure(pts * nout * 2) || c->nl_acc.ensure(pts * h.nentry * 2) ||
        c->nl_scr.ensure(std::max({(size_t) rhg::NLTE_NSLOT * (nlist - h.nbez),
                                   stokes_out ? (size_t) rhg::NLTE_NSLOT_EM * nlist : (size_t) 0,
                                   (size_t) 1}) *
                         Nmax * ncb_max * 64 * sizeof(double)) ||
        (stokes_out && !device_ptrs && c->nl_st.ensure((size_t) ncol * 4 * nout * sizeof(double))) ||
        c->nl_jpar.ensure(ncol * sizeof(int32_t)) ||
        c->nl_dJ.ensure((size_t) nlist * ncol * sizeof(double)) || c->nl_hist.ensure(pts * NS * nl) ||
        c->nl_act.ensure(ncol * sizeof(int32_t)) || c->nl_red.ensure(ncol * sizeof(double)) ||
        (!device_ptrs && (c->nl_atm.ensure(pts * 9) || c->nl_ndep.ensure(ncol * sizeof(int32_t)) ||
                          c->nl_out.ensure(pts * nl * (nstar_out ? 2 : 1)))))
      return fail(RHG_ENOMEM, "device allocation for the NLTE batch failed");
    const double *d_atm = (const double *) cols->atmos + (size_t) c0 * 9 * Nmax;
    const int32_t *d_ndep = cols->ndep + c0;
    if (!device_ptrs) {
      HIPCHECK(hipMemcpyAsync(c->nl_atm.p, d_atm, pts * 9, hipMemcpyHostToDevice, c->stream));
      HIPCHECK(hipMemcpyAsync(c->nl_ndep.p, d_ndep, ncol * sizeof(int32_t), hipMemcpyHostToDevice,
                              c->stream));
      d_atm = (const double *) c->nl_atm.p;
      d_ndep = (const int32_t *) c->nl_ndep.p;
    }
    rhg::NlteBufs B{};
    B.ncol = ncol;
    B.Nmax = Nmax;
    B.nfield = nfield;
    B.nout = nout;
    B.ndep = d_ndep;
    B.state = (double *) c->state.p;
    B.nstar = (double *) c->nl_nstar.p;
    B.pops = (double *) c->nl_pops.p;
    B.ntot = (double *) c->nl_ntot.p;
    B.vbroad = (double *) c->nl_vbroad.p;
    B.nHe = (double *) c->nl_nHe.p;
    B.adamp = (double *) c->nl_adamp.p;
    B.wphi = (double *) c->nl_wphi.p;
    B.Cmat = (double *) c->nl_Cmat.p;
    B.Gam = (double *) c->nl_Gam.p;
    B.phi = (double *) c->nl_phi.p;
    B.cchi = (double *) c->cchi.p;
    B.ceta = (double *) c->ceta.p;
    B.csca = (double *) c->nl_csca.p;
    B.J = (double *) c->nl_J.p;
    B.jpar = (int32_t *) c->nl_jpar.p;
    B.acc = (double *) c->nl_acc.p;
    B.scr = (double *) c->nl_scr.p;
    B.dJ = (double *) c->nl_dJ.p;
    B.hist = (double *) c->nl_hist.p;
    int32_t *d_act = (int32_t *) c->nl_act.p;
    double *d_red = (double *) c->nl_red.p;
    const int64_t npts = (int64_t) ncol * Nmax;
    const dim3 b64(64), gpts((unsigned) ((npts + 63) / 64));

    // set-up: column state, scattering in chi_ref, depth scale, background,
    // populations, collisions, profiles
    int rc = timed(0, [&]() -> int {
      if (int r = launch_state(c, d_atm, d_ndep, nullptr, ncol, Nmax, cols->mu, cols->atm_scale,
                               B.state, nfield))
        return r;
      hipLaunchKernelGGL(rhg::k_nlte_pre, gpts, b64, 0, c->stream, c->d_model, d_at, B);
      HIPCHECK(hipGetLastError());
      hipLaunchKernelGGL(rhg::k_geometry, dim3((unsigned) ((ncol + 63) / 64)), b64, 0, c->stream,
                         c->d_model, d_ndep, ncol, Nmax, cols->atm_scale, cols->mu, B.state, nfield,
                         (int32_t *) c->moving.p, (double *) nullptr);
      HIPCHECK(hipGetLastError());
      hipLaunchKernelGGL((rhg::k_continuum<double, false>),
                         dim3((unsigned) ((npts + rhg::CONT_BLOCK - 1) / rhg::CONT_BLOCK)),
                         dim3(rhg::CONT_BLOCK), 0, c->stream, c->d_model, d_ndep, ncol, Nmax,
                         B.state, nfield, c->d_out_ls, nout, B.cchi, B.ceta, (const int32_t *) nullptr, (const int32_t *) nullptr, (const int32_t *) nullptr, 0);
      HIPCHECK(hipGetLastError());
      hipLaunchKernelGGL(rhg::k_nlte_bg, gpts, b64, 0, c->stream, c->d_model, d_at, B,
                         c->d_out_ls);
      HIPCHECK(hipGetLastError());
      if (a->nline > 0) {
        const int64_t np = npts * a->nline;
        hipLaunchKernelGGL(rhg::k_nlte_profile, dim3((unsigned) ((np + 63) / 64)), b64, 0,
                           c->stream, c->d_model, d_at, B);
        HIPCHECK(hipGetLastError());
      }
      HIPCHECK(hipMemsetAsync(B.J, 0, pts * nout, c->stream));
      HIPCHECK(hipMemsetAsync(B.jpar, 0, ncol * sizeof(int32_t), c->stream));
      return 0;
    });
    if (rc) return rc;

    std::vector<int32_t> act(ncol), keep;
    std::vector<double> red(ncol);
    for (int i = 0; i < ncol; i++) act[i] = i;
    auto upload_act = [&]() -> int {
      HIPCHECK(hipMemcpyAsync(d_act, act.data(), act.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                              c->stream));
      return 0;
    };
    auto rt = [&](int slot, int eval) -> int {
      const int nact = (int) act.size();
      const int ncb = (nact + 63) / 64, nfe = nlist - h.nbez;
      int r = timed(slot, [&]() -> int {
        if (h.nbez > 0) {
          dim3 grd((unsigned) rhg::xcd_grid(ncb, h.nbez));
#define RHG_LAUNCH_BEZ(MS)                                                                       \
  hipLaunchKernelGGL((rhg::k_nlte_bezier<rhg::NLTE_RB, MS>), grd, b64, 0, c->stream, c->d_model, \
                     d_at, B, d_act, nact, c->d_out_ls, d_iolist, h.nbez, 0, eval)
          if (h.maxset_bez <= 2) RHG_LAUNCH_BEZ(2);
          else if (h.maxset_bez <= 4) RHG_LAUNCH_BEZ(4);
          else RHG_LAUNCH_BEZ(8);
#undef RHG_LAUNCH_BEZ
          HIPCHECK(hipGetLastError());
        }
        c->nlte_ms[6] += 1.0;
        c->nlte_ms[7] += (double) nact * nlist;
        c->nlte_ms[eval ? 11 : 10] += (double) nact;
        return 0;
      });
      if (r) return r;
      r = timed(8 + (slot == 2), [&]() -> int {
        if (nfe > 0) {
          // the member loops unrolled to the largest active set (one launch:
          // split by set size, the launches would not fill the GPU)
          dim3 grd((unsigned) rhg::xcd_grid(ncb, nfe));
#define RHG_LAUNCH_FE(MS)                                                                        \
  hipLaunchKernelGGL((rhg::k_nlte_feautrier<MS>), grd, b64, 0, c->stream, c->d_model, d_at, B,   \
                     d_act, nact, c->d_out_ls, d_iolist + h.nbez, nfe, h.nbez, eval)
          if (h.maxset_fe <= 2) RHG_LAUNCH_FE(2);
          else if (h.maxset_fe <= 4) RHG_LAUNCH_FE(4);
          else if (h.maxset_fe <= 6) RHG_LAUNCH_FE(6);
          else RHG_LAUNCH_FE(8);
#undef RHG_LAUNCH_FE
          HIPCHECK(hipGetLastError());
        }
        hipLaunchKernelGGL(rhg::k_nlte_flip, dim3((unsigned) ((nact + 63) / 64)), b64, 0, c->stream,
                           B, d_act, nact);
        HIPCHECK(hipGetLastError());
        return 0;
      });
      return r;
    };
    // per-column reduction of the active columns into red[]
    auto fetch_red = [&]() -> int {
      HIPCHECK(hipMemcpyAsync(red.data(), d_red, ncol * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHECK(hipStreamSynchronize(c->stream));
      return 0;
    };

    // background scattering passes without the operator (each column stops
    // when its own maximum change of J is below ITER_LIMIT)
    if (a->nmax_iter > 0) {
      if ((rc = upload_act())) return rc;
      for (int it = 0; it < a->nmax_scatter && !act.empty(); it++) {
        if ((rc = rt(1, 0))) return rc;
        const int nact = (int) act.size();
        hipLaunchKernelGGL(rhg::k_nlte_djmax, dim3((unsigned) ((nact + 63) / 64)), b64, 0, c->stream,
                           B, d_act, nact, nlist, d_red);
        HIPCHECK(hipGetLastError());
        if ((rc = fetch_red())) return rc;
        keep.clear();
        for (int ic : act)
          if (!(red[ic] < a->iter_limit)) keep.push_back(ic);
        act.swap(keep);
        if ((rc = upload_act())) return rc;
      }
    }

    // accelerated lambda iterations
    std::vector<int32_t> niter(ncol, 0);
    std::vector<double> dpops(ncol, 0.0);
    act.resize(ncol);
    for (int i = 0; i < ncol; i++) act[i] = i;
    if ((rc = upload_act())) return rc;
    HIPCHECK(hipMemcpyAsync(B.hist, B.pops, pts * nl, hipMemcpyDeviceToDevice, c->stream));
    const int ndelay = std::max(a->ng_delay, a->ng_order + 2);
    for (int it = 1; it <= a->nmax_iter && !act.empty(); it++) {
      const int nact = (int) act.size();
      if ((rc = rt(2, 1))) return rc;
      rc = timed(3, [&]() -> int {
        const int64_t n = (int64_t) nact * Nmax;
        const dim3 g((unsigned) ((n + 63) / 64));
        if (nl <= 4)
          hipLaunchKernelGGL(rhg::k_nlte_stateq<4>, g, b64, 0, c->stream, d_at, B, d_act, nact, a->isum);
        else if (nl <= 6)
          hipLaunchKernelGGL(rhg::k_nlte_stateq<6>, g, b64, 0, c->stream, d_at, B, d_act, nact, a->isum);
        else if (nl <= 8)
          hipLaunchKernelGGL(rhg::k_nlte_stateq<8>, g, b64, 0, c->stream, d_at, B, d_act, nact, a->isum);
        else
          hipLaunchKernelGGL(rhg::k_nlte_stateq<16>, g, b64, 0, c->stream, d_at, B, d_act, nact, a->isum);
        HIPCHECK(hipGetLastError());
        return 0;
      });
      if (rc) return rc;
      rc = timed(4, [&]() -> int {
        hipLaunchKernelGGL(rhg::k_nlte_ng, dim3((unsigned) ((nact + 63) / 64)), b64, 0, c->stream, d_at,
                           B, d_act, nact, it, a->ng_order, ndelay, a->ng_period, d_red);
        HIPCHECK(hipGetLastError());
        return 0;
      });
      if (rc) return rc;
      if ((rc = fetch_red())) return rc;
      keep.clear();
      for (int ic : act) {
        niter[ic] = it;
        dpops[ic] = red[ic];
        if (!(red[ic] < a->iter_limit)) keep.push_back(ic);
      }
      act.swap(keep);
      if (!act.empty() && (rc = upload_act())) return rc;
    }

    // the emergent spectrum: J with the converged populations (up to
    // N_MAX_SCATTER passes, each column stopping at max dJ <= ITER_LIMIT,
    // rh/rhf1d/pyrh_compute1dray.c:332-337), then the final ray
    if (stokes_out) {
      act.resize(ncol);
      for (int i = 0; i < ncol; i++) act[i] = i;
      if ((rc = upload_act())) return rc;
      for (int it = 0; it < a->nmax_scatter && !act.empty(); it++) {
        if ((rc = rt(1, 0))) return rc;
        const int nact = (int) act.size();
        hipLaunchKernelGGL(rhg::k_nlte_djmax, dim3((unsigned) ((nact + 63) / 64)), b64, 0, c->stream,
                           B, d_act, nact, nlist, d_red);
        HIPCHECK(hipGetLastError());
        if ((rc = fetch_red())) return rc;
        keep.clear();
        for (int ic : act)
          if (!(red[ic] <= a->iter_limit)) keep.push_back(ic);
        act.swap(keep);
        if (!act.empty() && (rc = upload_act())) return rc;
      }
      double *d_st = device_ptrs ? stokes_out + (size_t) c0 * 4 * nout : (double *) c->nl_st.p;
      dim3 grd((unsigned) rhg::xcd_grid((ncol + 63) / 64, nlist));
      hipLaunchKernelGGL(rhg::k_nlte_emergent, grd, b64, 0, c->stream, c->d_model, d_at, B,
                         c->d_out_ls, d_iolist, nlist, cols->mu, d_st);
      HIPCHECK(hipGetLastError());
      if (!device_ptrs)
        HIPCHECK(hipMemcpyAsync(stokes_out + (size_t) c0 * 4 * nout, d_st,
                                (size_t) ncol * 4 * nout * sizeof(double), hipMemcpyDeviceToHost,
                                c->stream));
    }

    // outputs of the chunk
    double *d_out = device_ptrs ? pops_out + (size_t) c0 * nl * Nmax : (double *) c->nl_out.p;
    double *d_nst = !nstar_out ? nullptr
                    : device_ptrs ? nstar_out + (size_t) c0 * nl * Nmax
                                  : (double *) c->nl_out.p + (size_t) ncol * nl * Nmax;
    {
      const int64_t n = (int64_t) ncol * nl * Nmax;
      hipLaunchKernelGGL(rhg::k_nlte_out, dim3((unsigned) ((n + 63) / 64)), b64, 0, c->stream, d_at, B,
                         (const double *) B.pops, d_out);
      HIPCHECK(hipGetLastError());
      if (d_nst) {
        hipLaunchKernelGGL(rhg::k_nlte_out, dim3((unsigned) ((n + 63) / 64)), b64, 0, c->stream, d_at,
                           B, (const double *) B.nstar, d_nst);
        HIPCHECK(hipGetLastError());
        if (!device_ptrs)
          HIPCHECK(hipMemcpyAsync(nstar_out + (size_t) c0 * nl * Nmax, d_nst, pts * nl,
                                  hipMemcpyDeviceToHost, c->stream));
      }
    }
    if (device_ptrs) {
      HIPCHECK(hipMemcpyAsync(niter_out + c0, niter.data(), ncol * sizeof(int32_t), hipMemcpyHostToDevice,
                              c->stream));
      HIPCHECK(hipMemcpyAsync(dpops_out + c0, dpops.data(), ncol * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
    } else {
      HIPCHECK(hipMemcpyAsync(pops_out + (size_t) c0 * nl * Nmax, d_out, pts * nl, hipMemcpyDeviceToHost,
                              c->stream));
      memcpy(niter_out + c0, niter.data(), ncol * sizeof(int32_t));
      memcpy(dpops_out + c0, dpops.data(), ncol * sizeof(double));
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
  }
  HIPCHECK(hipEventRecord(tm_all.e[1], c->stream));
  HIPCHECK(hipEventSynchronize(tm_all.e[1]));
  HIPCHECK(hipEventElapsedTime(&ms, tm_all.e[0], tm_all.e[1]));
  c->nlte_ms[5] = ms;
  g_err.clear();
  return 0;
}

// Validates the element tables of an rhg_hse_t, uploads them on the
// context's stream and fills H (without its scratch).
static int hse_tables(rhg_context_t *c, const rhg_hse_t *h, rhg::DevHse &H) {
  const rhg::DevModel &dm = c->dm;
  if (h->nel < 1 || !h->el_nstage || !h->el_pf_ptr || !h->el_logpf || !h->el_ionpot || !h->el_abund)
    return fail(RHG_EINVAL, "empty element table");
  if (h->lam_index < 0 || h->lam_index >= c->nspect) return fail(RHG_EINVAL, "lam_index out of range");
  if (h->he_atom >= dm.natom) return fail(RHG_EINVAL, "he_atom out of range");
  if (h->nmax_iter < 1) return fail(RHG_EINVAL, "nmax_iter must be positive");
  // the element tables are packed: element e's stages start where e-1's end
  int64_t npfel = 0;
  for (int e = 0; e < h->nel; e++) {
    if (h->el_nstage[e] < 1 || h->el_nstage[e] > rhg::NSTAGE_MAX)
      return fail(RHG_EUNSUPPORTED, "element with too many ionisation stages");
    if (h->el_pf_ptr[e] != npfel) return fail(RHG_EINVAL, "element tables are not packed");
    npfel += (int64_t) h->el_nstage[e] * dm.npf;
  }
  if (c->hs_nstage.ensure(h->nel * sizeof(int32_t)) || c->hs_ptr.ensure(h->nel * sizeof(int32_t)) ||
      c->hs_logpf.ensure(npfel * sizeof(double)) || c->hs_ionpot.ensure(npfel * sizeof(double)) ||
      c->hs_abund.ensure(h->nel * sizeof(double)))
    return fail(RHG_ENOMEM, "device allocation for the element tables failed");
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipMemcpyAsync(c->hs_nstage.p, h->el_nstage, h->nel * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->hs_ptr.p, h->el_pf_ptr, h->nel * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->hs_logpf.p, h->el_logpf, npfel * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->hs_ionpot.p, h->el_ionpot, npfel * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->hs_abund.p, h->el_abund, h->nel * sizeof(double), hipMemcpyHostToDevice, c->stream));
  H.nel = h->nel;
  H.el_nstage = (const int32_t *) c->hs_nstage.p;
  H.el_pf_ptr = (const int32_t *) c->hs_ptr.p;
  H.el_logpf = (const double *) c->hs_logpf.p;
  H.el_ionpot = (const double *) c->hs_ionpot.p;
  H.el_abund = (const double *) c->hs_abund.p;
  H.lam_index = h->lam_index;
  H.he_atom = h->he_atom;
  H.ray_H = h->ray_H;
  H.ray_He = h->ray_He;
  H.ray_H2 = h->ray_H2;
  H.sigma_thomson = h->sigma_thomson;
  H.hmin_fudge = h->hmin_fudge;
  H.scatt_fudge = h->scatt_fudge;
  H.nmax_iter = h->nmax_iter;
  H.iter_limit = h->iter_limit;
  H.scr = nullptr;
  H.nslot_stage = 0;
  return 0;
}

int rhg_hse(rhg_context_t *c, const rhg_hse_t *h, int32_t ncol, int32_t Nmax, const int32_t *ndep,
            int32_t atm_scale, const double *scale, const double *temp, const double *pg_top,
            int32_t device_ptrs, double *ne_out, double *nHtot_out, double *rho_out,
            double *pg_out, int32_t *niter_out) {
  if (!c || !h || !ndep || !scale || !temp || !pg_top || !ne_out || !nHtot_out || !rho_out ||
      !pg_out)
    return fail(RHG_EINVAL, "null argument");
  if (ncol < 1 || Nmax < 1 || Nmax > RHG_MAX_DEPTH) return fail(RHG_EINVAL, "bad batch shape");
  if ((int64_t) ncol * Nmax > (int64_t) INT32_MAX / 16) return fail(RHG_EINVAL, "batch too large");
  if (atm_scale == 1)
    return fail(RHG_EUNSUPPORTED, "hse on the column-mass scale: the reference leaves the "
                                  "pressure step undefined there (pyrh_hse.c:282-290)");
  if (atm_scale != 0 && atm_scale != 2) return fail(RHG_EINVAL, "bad atm_scale");
  const rhg::DevModel &dm = c->dm;
  if (dm.nnuclei < 1 || dm.nnuclei > 8) return fail(RHG_EUNSUPPORTED, "number of molecule nuclei out of range");
  if (dm.natom < 1 || c->h_nucl_atom.empty() || c->h_nucl_atom[0] != 0)
    return fail(RHG_EINVAL, "hydrogen must be the first atom and the first nucleus");
  std::vector<int32_t> hndep(ncol);
  if (device_ptrs) {
    HIPCHECK(hipMemcpy(hndep.data(), ndep, ncol * sizeof(int32_t), hipMemcpyDeviceToHost));
  } else {
    memcpy(hndep.data(), ndep, ncol * sizeof(int32_t));
  }
  for (int i = 0; i < ncol; i++)
    if (hndep[i] < 1 || hndep[i] > Nmax) return fail(RHG_EINVAL, "ndep out of range");
  HIPCHECK(hipSetDevice(c->device));

  const size_t col_b = (size_t) ncol * Nmax * sizeof(double);
  rhg::DevHse H;
  if (int rc = hse_tables(c, h, H)) return rc;
  if (!device_ptrs && (c->hs_in.ensure(2 * col_b + ncol * sizeof(double)) ||
                       c->hs_ndep.ensure(ncol * sizeof(int32_t)) || c->hs_out.ensure(4 * col_b) ||
                       (niter_out && c->hs_niter.ensure((size_t) ncol * Nmax * sizeof(int32_t)))))
    return fail(RHG_ENOMEM, "device allocation for the batch failed");
  const double *d_scale = scale, *d_temp = temp, *d_pgt = pg_top;
  const int32_t *d_ndep = ndep;
  double *d_ne = ne_out, *d_nH = nHtot_out, *d_rho = rho_out, *d_pg = pg_out;
  int32_t *d_niter = niter_out;
  if (!device_ptrs) {
    double *in = (double *) c->hs_in.p;
    HIPCHECK(hipMemcpyAsync(in, scale, col_b, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(in + (size_t) ncol * Nmax, temp, col_b, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(in + (size_t) 2 * ncol * Nmax, pg_top, ncol * sizeof(double),
                            hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->hs_ndep.p, ndep, ncol * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    d_scale = in;
    d_temp = in + (size_t) ncol * Nmax;
    d_pgt = in + (size_t) 2 * ncol * Nmax;
    d_ndep = (const int32_t *) c->hs_ndep.p;
    double *out = (double *) c->hs_out.p;
    d_ne = out;
    d_nH = out + (size_t) ncol * Nmax;
    d_rho = out + (size_t) 2 * ncol * Nmax;
    d_pg = out + (size_t) 3 * ncol * Nmax;
    d_niter = niter_out ? (int32_t *) c->hs_niter.p : nullptr;
  }
  // lanes per column: enough lane groups to give every SIMD of the chip two
  // waves (256 CUs x 4 SIMDs x 64 lanes x 2, the occupancy k_hse is built
  // for), at most 16
  int lpc = 1;
  while (lpc < 16 && (int64_t) ncol * lpc < 131072) lpc <<= 1;
  if (const char *lanes_env = getenv("RHG_HSE_LANES")) {
    const int v = atoi(lanes_env);
    if (v == 1 || v == 2 || v == 4 || v == 8 || v == 16) lpc = v;
  }
  c->hse_lpc = lpc;
  const int64_t nthr = (int64_t) ncol * lpc;
  dim3 blk(rhg::HSE_BLOCK), grd((unsigned) ((nthr + rhg::HSE_BLOCK - 1) / rhg::HSE_BLOCK));
  // stage-factor slots: the most stages (beyond the first) any group lane
  // handles, its elements being sub, sub + lpc, ...
  int nslot = 1;
  for (int sub = 0; sub < lpc; sub++) {
    int n = 0;
    for (int e = sub; e < h->nel; e += lpc) n += h->el_nstage[e] - 1;
    nslot = std::max(nslot, n);
  }
  // the level groups: every molecule nucleus and He, ionisation differences
  // below HSE_NDZ
  for (int i = 0; i <= dm.nnuclei; i++) {
    const int ia = i < dm.nnuclei ? c->h_nucl_atom[i] : h->he_atom;
    if (ia < 0) continue;
    const int l0 = c->h_atom_level_ptr[ia], l1 = c->h_atom_level_ptr[ia + 1];
    for (int l = l0; l < l1; l++)
      if (c->h_level_stage[l] - c->h_level_stage[l0] >= rhg::HSE_NDZ)
        return fail(RHG_EUNSUPPORTED, "model atom with levels four or more ionisation stages "
                                      "above its ground level");
  }
  const int64_t nlaunch = (int64_t) grd.x * rhg::HSE_BLOCK;
  const int64_t nslot_all = nslot + (int64_t) (dm.nnuclei + 1) * (rhg::HSE_NDZ + 1);
  if (c->hs_scr.ensure((size_t) nslot_all * nlaunch * sizeof(double)))
    return fail(RHG_ENOMEM, "device allocation for the layer tables failed");
  H.scr = (double *) c->hs_scr.p;
  H.nslot_stage = nslot;
  const size_t lds = (size_t) 3 * dm.nmol * rhg::HSE_BLOCK * sizeof(double);
  HIPCHECK(hipEventRecord(c->ev[0], c->stream));
  switch (dm.nnuclei) {
#define RHG_LAUNCH_HSE(NN)                                                                          \
  case NN:                                                                                          \
    hipLaunchKernelGGL(rhg::k_hse<NN>, grd, blk, lds, c->stream, c->d_model, H, lpc, ncol, Nmax,    \
                       d_ndep, atm_scale, d_scale, d_temp, d_pgt, d_ne, d_nH, d_rho, d_pg, d_niter); \
    break;
    RHG_LAUNCH_HSE(1) RHG_LAUNCH_HSE(2) RHG_LAUNCH_HSE(3) RHG_LAUNCH_HSE(4)
    RHG_LAUNCH_HSE(5) RHG_LAUNCH_HSE(6) RHG_LAUNCH_HSE(7) RHG_LAUNCH_HSE(8)
#undef RHG_LAUNCH_HSE
  }
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipEventRecord(c->ev[1], c->stream));
  if (!device_ptrs) {
    HIPCHECK(hipMemcpyAsync(ne_out, d_ne, col_b, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipMemcpyAsync(nHtot_out, d_nH, col_b, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipMemcpyAsync(rho_out, d_rho, col_b, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipMemcpyAsync(pg_out, d_pg, col_b, hipMemcpyDeviceToHost, c->stream));
    if (niter_out)
      HIPCHECK(hipMemcpyAsync(niter_out, d_niter, (size_t) ncol * Nmax * sizeof(int32_t),
                              hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipEventElapsedTime(&c->hse_ms, c->ev[0], c->ev[1]));
  return 0;
}

int rhg_ne_from_nh(rhg_context_t *c, const rhg_hse_t *h, int32_t ncol, int32_t Nmax,
                   const int32_t *ndep, const double *temp, const double *nH,
                   int32_t device_ptrs, double *ne_out) {
  if (!c || !h || !ndep || !temp || !nH || !ne_out) return fail(RHG_EINVAL, "null argument");
  if (ncol < 1 || Nmax < 1 || Nmax > RHG_MAX_DEPTH) return fail(RHG_EINVAL, "bad batch shape");
  if ((int64_t) ncol * Nmax > (int64_t) INT32_MAX / 16) return fail(RHG_EINVAL, "batch too large");
  std::vector<int32_t> hndep(ncol);
  if (device_ptrs) {
    HIPCHECK(hipMemcpy(hndep.data(), ndep, ncol * sizeof(int32_t), hipMemcpyDeviceToHost));
  } else {
    memcpy(hndep.data(), ndep, ncol * sizeof(int32_t));
  }
  for (int i = 0; i < ncol; i++)
    if (hndep[i] < 1 || hndep[i] > Nmax) return fail(RHG_EINVAL, "ndep out of range");
  HIPCHECK(hipSetDevice(c->device));
  rhg::DevHse H;
  if (int rc = hse_tables(c, h, H)) return rc;
  const size_t col_b = (size_t) ncol * Nmax * sizeof(double);
  const double *d_T = temp, *d_nH = nH;
  const int32_t *d_ndep = ndep;
  double *d_ne = ne_out;
  if (!device_ptrs) {
    if (c->hs_in.ensure(2 * col_b) || c->hs_ndep.ensure(ncol * sizeof(int32_t)) ||
        c->hs_out.ensure(col_b))
      return fail(RHG_ENOMEM, "device allocation for the batch failed");
    double *in = (double *) c->hs_in.p;
    HIPCHECK(hipMemcpyAsync(in, temp, col_b, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(in + (size_t) ncol * Nmax, nH, col_b, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->hs_ndep.p, ndep, ncol * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    d_T = in;
    d_nH = in + (size_t) ncol * Nmax;
    d_ndep = (const int32_t *) c->hs_ndep.p;
    d_ne = (double *) c->hs_out.p;
  }
  const int64_t npts = (int64_t) ncol * Nmax;
  dim3 blk(rhg::HSE_BLOCK), grd((unsigned) ((npts + rhg::HSE_BLOCK - 1) / rhg::HSE_BLOCK));
  hipLaunchKernelGGL(rhg::k_ne_from_nh, grd, blk, 0, c->stream, c->d_model, H, ncol, Nmax, d_ndep,
                     d_T, d_nH, d_ne);
  HIPCHECK(hipGetLastError());
  if (!device_ptrs) HIPCHECK(hipMemcpyAsync(ne_out, d_ne, col_b, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int rhg_scales(rhg_context_t *c, const rhg_columns_t *cols, int32_t device_ptrs, double *tau_out,
               double *height_out, double *cmass_out) {
  if (!c || !cols || !tau_out || !height_out || !cmass_out) return fail(RHG_EINVAL, "null argument");
  const int ncol = cols->ncol, Nmax = cols->ndep_max;
  if (ncol < 1 || Nmax < 2 || Nmax > RHG_MAX_DEPTH) return fail(RHG_EINVAL, "bad batch shape");
  if ((int64_t) ncol * Nmax > (int64_t) INT32_MAX / 16) return fail(RHG_EINVAL, "batch too large");
  if (cols->atm_scale < 0 || cols->atm_scale > 2) return fail(RHG_EINVAL, "bad atm_scale");
  if (c->ref_has_line)
    return fail(RHG_EUNSUPPORTED, "a Kurucz line at the reference wavelength: its opacity in "
                                  "the depth scales depends on the ray in moving atmospheres");
  if (c->dm.nnuclei < 1 || c->dm.nnuclei > 8)
    return fail(RHG_EUNSUPPORTED, "number of molecule nuclei out of range");
  std::vector<int32_t> hndep(ncol);
  if (device_ptrs) {
    HIPCHECK(hipMemcpy(hndep.data(), cols->ndep, ncol * sizeof(int32_t), hipMemcpyDeviceToHost));
  } else {
    memcpy(hndep.data(), cols->ndep, ncol * sizeof(int32_t));
  }
  for (int i = 0; i < ncol; i++)
    if (hndep[i] < 2 || hndep[i] > Nmax) return fail(RHG_EINVAL, "ndep out of range");
  HIPCHECK(hipSetDevice(c->device));
  const int nfield = rhg::state_nfield(c->ncont, c->nline);
  const size_t atm_b = (size_t) ncol * 9 * Nmax * sizeof(double);
  const size_t col_b = (size_t) ncol * Nmax * sizeof(double);
  if (c->state.ensure((size_t) ncol * nfield * Nmax * sizeof(double)) ||
      (!device_ptrs && (c->atmos.ensure(atm_b) || c->ndep.ensure(ncol * sizeof(int32_t)) ||
                        c->hs_out.ensure(3 * col_b))))
    return fail(RHG_ENOMEM, "device allocation for the batch failed");
  const double *d_atm = cols->atmos;
  const int32_t *d_ndep = cols->ndep;
  double *d_tau = tau_out, *d_h = height_out, *d_m = cmass_out;
  if (!device_ptrs) {
    HIPCHECK(hipMemcpyAsync(c->atmos.p, cols->atmos, atm_b, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->ndep.p, cols->ndep, ncol * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    d_atm = (const double *) c->atmos.p;
    d_ndep = (const int32_t *) c->ndep.p;
    d_tau = (double *) c->hs_out.p;
    d_h = d_tau + (size_t) ncol * Nmax;
    d_m = d_h + (size_t) ncol * Nmax;
  }
  double *state = (double *) c->state.p;
  if (int rc = launch_state(c, d_atm, d_ndep, nullptr, ncol, Nmax, 1.0, cols->atm_scale, state, nfield))
    return rc;
  dim3 blk(64), grd((unsigned) ((ncol + 63) / 64));
  hipLaunchKernelGGL(rhg::k_scales, grd, blk, 0, c->stream, c->d_model, d_ndep, ncol, Nmax,
                     cols->atm_scale, state, nfield, d_tau, d_h, d_m);
  HIPCHECK(hipGetLastError());
  if (!device_ptrs) {
    HIPCHECK(hipMemcpyAsync(tau_out, d_tau, col_b, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipMemcpyAsync(height_out, d_h, col_b, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipMemcpyAsync(cmass_out, d_m, col_b, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int rhg_hse_timing(rhg_context_t *ctx, double *ms_out, int32_t *lanes_out) {
  if (!ctx) return fail(RHG_EINVAL, "null argument");
  if (ms_out) *ms_out = ctx->hse_ms;
  if (lanes_out) *lanes_out = ctx->hse_lpc;
  return 0;
}

int rhg_nlte_timings(rhg_context_t *ctx, double *out) {
  if (!ctx || !out) return fail(RHG_EINVAL, "null argument");
  for (int s = 0; s < RHG_NLTE_NTIMING; s++) out[s] = ctx->nlte_ms[s];
  return 0;
}

int rhg_set_tuning(rhg_context_t *ctx, int32_t key, int32_t value) {
  if (!ctx) return fail(RHG_EINVAL, "null argument");
  if (value != 0 && value != 1) return fail(RHG_EINVAL, "tuning values are 0 or 1");
  switch (key) {
    case RHG_TUNE_LINE_TABLE: ctx->tune_line_table = value; break;
    case RHG_TUNE_COLUMN_SORT: ctx->tune_column_sort = value; break;
    case RHG_TUNE_WAVE_ORDER: {
      const std::vector<int32_t> &v = value ? ctx->h_pol_heavy : ctx->h_pol;
      if (!v.empty()) {
        HIPCHECK(hipSetDevice(ctx->device));
        HIPCHECK(hipMemcpy((void *) ctx->d_io_pol, v.data(), v.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice));
      }
      ctx->tune_wave_order = value;
      break;
    }
    default: return fail(RHG_EINVAL, "unknown tuning key");
  }
  g_err.clear();
  return 0;
}

int rhg_get_tuning(rhg_context_t *ctx, int32_t key, int32_t *value) {
  if (!ctx || !value) return fail(RHG_EINVAL, "null argument");
  switch (key) {
    case RHG_TUNE_LINE_TABLE: *value = ctx->tab_bytes > 0 && ctx->tune_line_table; break;
    case RHG_TUNE_COLUMN_SORT: *value = ctx->tune_column_sort; break;
    case RHG_TUNE_WAVE_ORDER: *value = ctx->tune_wave_order; break;
    default: return fail(RHG_EINVAL, "unknown tuning key");
  }
  return 0;
}

int rhg_last_timings(rhg_context_t *ctx, double *ms_out) {
  if (!ctx || !ms_out) return fail(RHG_EINVAL, "null argument");
  for (int s = 0; s < 4; s++) ms_out[s] = ctx->last_ms[s];
  return 0;
}

int rhg_rf_timings(rhg_context_t *ctx, double *ms_out) {
  if (!ctx || !ms_out) return fail(RHG_EINVAL, "null argument");
  for (int s = 0; s < 5; s++) ms_out[s] = ctx->rf_ms[s];
  return 0;
}

}  // extern "C"
