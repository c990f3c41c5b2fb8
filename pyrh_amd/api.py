"""pyrh-compatible Python API on top of the GPU engine.

``compute1d`` keeps the signature and return tuple of pyrh.compute1d
(pyrh.pyx:537-668): LTE runs go through the LTE synthesis, runs with an
ACTIVE atom through the NLTE solve (populations, then the emergent ray);
``synth_batch`` and ``nlte_batch`` are the batched forms that the GPU is
built for.  ``hse``, ``get_ne_from_nH`` and ``get_scales`` keep the
signatures of pyrh's (pyrh.pyx:396-533); ``hse_batch`` is hse's batched
form.  Parsed inputs, models and device contexts are
cached per run directory and wavelength grid, so repeated calls do not
re-read files (the reference re-reads everything on every call,
rh/rhf1d/pyrh_compute1dray.c:130-301).
"""
from __future__ import annotations

import hashlib
import os

import numpy as np

from . import hse as _hse
from . import model as _model
from . import nlte as _nlte
from .engine import Engine

_SETUPS = {}
_ENGINES = {}
_NLTE = {}
_HSE = {}


class Populations:
    """Populations of one active atom, pyrh.Populations (pyrh.pyx:170-176):
    ID, nlevel, nz (depth points), n and nstar [nlevel, nz]."""

    def __init__(self, ID, nlevel, nz, n, nstar):
        self.ID = ID
        self.nlevel = nlevel
        self.nz = nz
        self.n = n
        self.nstar = nstar


def _key(*arrays):
    h = hashlib.sha1()
    for a in arrays:
        if a is None:
            h.update(b"-")
        else:
            a = np.ascontiguousarray(a)
            h.update(str(a.dtype).encode() + str(a.shape).encode())
            h.update(a.tobytes())
    return h.hexdigest()


def _setup(cwd, atomic_number, atomic_abundance, loggf_ids, loggf_values, lam_ids, lam_values):
    key = (os.path.abspath(cwd), os.environ.get("PYRH_PATH", ""),
           _key(atomic_number, atomic_abundance, loggf_ids, loggf_values, lam_ids, lam_values))
    s = _SETUPS.get(key)
    if s is None:
        s = _model.open_run(cwd, atomic_number=atomic_number, atomic_abundance=atomic_abundance,
                            loggf_ids=loggf_ids, loggf_values=loggf_values,
                            lam_ids=lam_ids, lam_values=lam_values)
        _SETUPS[key] = s
    return key, s


def _fudge(fudge_wave, fudge_value):
    """pyrh's fudge_wave/fudge_value as one (wave, value) pair, or None."""
    if fudge_wave is None and fudge_value is None:
        return None
    if fudge_wave is None or fudge_value is None:
        raise ValueError("fudge_wave and fudge_value go together")
    return (np.ascontiguousarray(fudge_wave, np.float64),
            np.ascontiguousarray(fudge_value, np.float64))


def engine_for(cwd, wave, atomic_number=None, atomic_abundance=None, loggf_ids=None,
               loggf_values=None, lam_ids=None, lam_values=None, device=0, fudge_wave=None,
               fudge_value=None):
    skey, s = _setup(cwd, atomic_number, atomic_abundance, loggf_ids, loggf_values,
                     lam_ids, lam_values)
    fudge = _fudge(fudge_wave, fudge_value)
    key = (skey, _key(np.asarray(wave, np.float64)), device,
           _key(*(fudge if fudge is not None else (None,))))
    e = _ENGINES.get(key)
    if e is None:
        m = _model.build_model(s, np.asarray(wave, np.float64), fudge=fudge)
        e = Engine(m, device)
        _ENGINES[key] = e
    return e


def nlte_engine_for(cwd, wave, atomic_number=None, atomic_abundance=None, loggf_ids=None,
                    loggf_values=None, lam_ids=None, lam_values=None, device=0,
                    fudge_wave=None, fudge_value=None):
    """(engine, NLTE model) for a run directory with one ACTIVE atom."""
    skey, s = _setup(cwd, atomic_number, atomic_abundance, loggf_ids, loggf_values,
                     lam_ids, lam_values)
    fudge = _fudge(fudge_wave, fudge_value)
    key = (skey, _key(np.asarray(wave, np.float64)), device,
           _key(*(fudge if fudge is not None else (None,))))
    r = _NLTE.get(key)
    if r is None:
        nm = _nlte.build_nlte(s, np.asarray(wave, np.float64), fudge=fudge)
        r = (Engine(nm.lte, device), nm)
        _NLTE[key] = r
    return r


def hse_engine_for(cwd, fudge_wave=None, fudge_value=None, atomic_number=None,
                   atomic_abundance=None, device=0):
    """(engine, HSE model) of pyrh's hse() for a run directory."""
    skey, s = _setup(cwd, atomic_number, atomic_abundance, None, None, None, None)
    key = (skey, _key(fudge_wave, fudge_value), device)
    r = _HSE.get(key)
    if r is None:
        hm = _hse.build_hse(s, fudge_wave, fudge_value)
        r = (Engine(hm.model, device), hm)
        _HSE[key] = r
    return r


def hse(cwd, atm_scale, scale, temp, pg_top=0.1, fudge_wave=None, fudge_value=None,
        atomic_number=None, atomic_abundance=None, full_output=False):
    """pyrh.hse (pyrh.pyx:427-487): gas pressure in hydrostatic equilibrium
    for one temperature column on `scale` (log10 tau500 for atm_scale 0,
    height [km] for 2), starting from pg_top [Pa] at the top.  Returns
    (ne, nHtot) in m^-3, or (ne, nHtot, rho, pg) in SI units with
    full_output."""
    scale = np.ascontiguousarray(scale, np.float64)
    temp = np.ascontiguousarray(temp, np.float64)
    if scale.ndim != 1 or temp.shape != scale.shape:
        raise ValueError("scale and temp must be 1-D arrays of the same length")
    _hse.check_scale(atm_scale)
    e, hm = hse_engine_for(cwd, fudge_wave, fudge_value, atomic_number, atomic_abundance)
    ne, nH, rho, pg, _ = e.hse(hm, scale[None], temp[None], pg_top, atm_scale)
    if full_output:
        return ne[0], nH[0], rho[0], pg[0]
    return ne[0], nH[0]


def hse_batch(cwd, atm_scale, scale, temp, pg_top=0.1, ndep=None, fudge_wave=None,
              fudge_value=None, atomic_number=None, atomic_abundance=None, device=0):
    """hse for a batch of columns: temp [ncol, ndep_max], scale the same
    shape or one row for all, pg_top scalar or [ncol], ndep [ncol] the depth
    points each column uses.  A dict with ne, nHtot, rho, pg [ncol,
    ndep_max] (SI units, zero past ndep) and the pressure iterations per
    layer."""
    _hse.check_scale(atm_scale)
    e, hm = hse_engine_for(cwd, fudge_wave, fudge_value, atomic_number, atomic_abundance, device)
    ne, nH, rho, pg, it = e.hse(hm, scale, temp, pg_top, atm_scale, ndep)
    return {"ne": ne, "nHtot": nH, "rho": rho, "pg": pg, "niter": it}


def get_ne_from_nH(cwd, atm_scale, scale, temperature, nH):
    """pyrh.get_ne_from_nH (pyrh.pyx:396-423, rh/rhf1d/pyrh_hse.c:555-678):
    the electron density [cm^-3] of LTE ionisation at every depth from the
    temperature [K] and n_Htot [cm^-3].  The depth scale does not enter
    (the reference sets it up and does not use it)."""
    temperature = np.ascontiguousarray(temperature, np.float64)
    nH = np.ascontiguousarray(nH, np.float64)
    if temperature.ndim != 1 or nH.shape != temperature.shape or np.size(scale) != nH.size:
        raise ValueError("scale, temperature and nH must be 1-D arrays of the same length")
    e, hm = hse_engine_for(cwd)
    return e.ne_from_nh(hm, temperature[None], nH[None])[0]


_SCALES = {}


def scales_engine_for(cwd, lam_ref, atomic_number=None, atomic_abundance=None, device=0):
    """Engine for pyrh's get_scales: the background model of the run
    directory on the single wavelength lam_ref, which is also its reference
    wavelength (pyrh_hse.c:475-479; Kurucz lines are read,
    rh/background.c:288)."""
    skey, s = _setup(cwd, atomic_number, atomic_abundance, None, None, None, None)
    key = (skey, float(lam_ref), device)
    e = _SCALES.get(key)
    if e is None:
        if any(a.active for a in s.atoms):
            raise NotImplementedError("get_scales with an ACTIVE atom in the atom list")
        m = _model.build_model(s, np.array([float(lam_ref)]), check=False, lam_ref=lam_ref)
        e = Engine(m, device)
        _SCALES[key] = e
    return e


def get_scales(cwd, atm_scale, scale, atmosphere, lam_ref, atomic_number=None,
               atomic_abundance=None):
    """pyrh.get_scales (pyrh.pyx:491-533, rh/rhf1d/pyrh_hse.c:402-553):
    tau (at lam_ref), height [m] and column mass [kg m^-2] of an atmosphere
    (rows as compute1d's; T, n_e, v_LOS, v_mic and n_H are read) given on
    `scale`.  The input scale is returned converted to SI units, where the
    reference leaves that array unset."""
    atm = np.array(atmosphere, np.float64)
    if atm.ndim != 2 or atm.shape[0] < 9:
        raise ValueError("atmosphere must be [>=9, ndep]")
    atm = np.ascontiguousarray(atm[:9])
    atm[0] = np.asarray(scale, np.float64)
    e = scales_engine_for(cwd, lam_ref, atomic_number, atomic_abundance)
    tau, h, m = e.scales(atm[None], atm_scale)
    return tau[0], h[0], m[0]


def has_active_atom(cwd):
    s = _setup(cwd, None, None, None, None, None, None)[1]
    return any(a.active for a in s.atoms)


def nlte_batch(cwd, mu, atm_scale, atmospheres, wave, ndep=None, device=0):
    """NLTE solve of a batch of columns [ncol, >=9, ndep_max] for a run
    directory with one ACTIVE atom: a dict with the populations n and nstar
    [ncol, nlevel, ndep_max], the iterations and last maximum relative change
    per column, the emergent Stokes I, Q, U, V [ncol, nlw] at mu and the
    wavelengths [nlw]."""
    e, nm = nlte_engine_for(cwd, wave, device=device)
    atm = np.asarray(atmospheres, np.float64)
    pops, niter, dpops, st, nst = e.nlte_solve(nm, atm, mu, atm_scale, ndep, stokes=True,
                                               nstar=True)
    lam = e.model.lam[e.model.out_index]
    return {"n": pops, "nstar": nst, "niter": niter, "dpops": dpops, "I": st[:, 0],
            "Q": st[:, 1], "U": st[:, 2], "V": st[:, 3], "lam": lam}


def synth_batch(cwd, mu, atm_scale, atmospheres, wave, ndep=None, precision="fp64",
                atomic_number=None, atomic_abundance=None, device=0, fudge_wave=None,
                fudge_value=None):
    """Stokes I, Q, U, V [ncol, nlw] and the wavelengths [nlw] for a batch
    of columns [ncol, >=9, ndep_max] (pyrh row order and units)."""
    e = engine_for(cwd, wave, atomic_number, atomic_abundance, device=device,
                   fudge_wave=fudge_wave, fudge_value=fudge_value)
    prec = {"fp64": 0, "fp32": 1}[precision]
    st = e.synth_lte(np.asarray(atmospheres, np.float64), mu, atm_scale, ndep, prec)
    lam = e.model.lam[e.model.out_index]
    return st[:, 0], st[:, 1], st[:, 2], st[:, 3], lam


def grid_positions(model, waves):
    """Positions of each wavelength grid's outputs on model's output grid:
    for grid g, the output wavelengths compute1d(..., waves[g]) returns
    (SortLambda: sorted, duplicates and LAMBDA_REF dropped,
    rh/sortlambda.c:42-225), located on the merged model."""
    out_lam = model.lam[model.out_index]
    lam_ref = model.lam[model.ref_index] if model.ref_index >= 0 else None
    pos = []
    for w in waves:
        u = np.unique(np.asarray(w, np.float64).ravel())
        if lam_ref is not None:
            u = u[u != lam_ref]
        p = np.searchsorted(out_lam, u)
        if np.any(p >= out_lam.size) or np.any(out_lam[np.minimum(p, out_lam.size - 1)] != u):
            raise ValueError("a grid wavelength is not on the merged grid")
        pos.append(p.astype(np.int64))
    return pos


def synth_grids(cwd, mu, atm_scale, atmospheres, waves, grid_of, ndep=None, precision="fp64",
                atomic_number=None, atomic_abundance=None, device=0):
    """LTE synthesis of a batch whose columns ask for different wavelength
    grids (column i for waves[grid_of[i]]), in one launch per stage.
    Returns I, Q, U, V [ncol, nlw] on the merged grid lam [nlw] (zero
    outside a column's grid) and pos, the list of each grid's positions on
    it: column i's spectrum on its own grid is I[i, pos[grid_of[i]]], equal
    to synth_batch with wave = waves[grid_of[i]]."""
    if len(waves) < 1:
        raise ValueError("waves must hold at least one grid")
    merged = np.unique(np.concatenate([np.asarray(w, np.float64).ravel() for w in waves]))
    e = engine_for(cwd, merged, atomic_number, atomic_abundance, device=device)
    pos = grid_positions(e.model, waves)
    if getattr(e, "grids", None) is None or len(e.grids) != len(pos) or \
            any(a.size != b.size or np.any(a != b) for a, b in zip(e.grids, pos)):
        e.set_grids(pos)
    prec = {"fp64": 0, "fp32": 1}[precision]
    st = e.synth_lte_grids(np.asarray(atmospheres, np.float64), grid_of, mu, atm_scale, ndep, prec)
    lam = e.model.lam[e.model.out_index]
    return st[:, 0], st[:, 1], st[:, 2], st[:, 3], lam, pos


def compute1d(cwd, mu, atm_scale, atmosphere, wave, loggf_ids=None, loggf_values=None,
              lam_ids=None, lam_values=None, fudge_wave=None, fudge_value=None,
              atomic_number=None, atomic_abundance=None, get_atomic_rfs=False,
              get_populations=False):
    """pyrh.compute1d (pyrh.pyx:537-668): returns (I, Q, U, V, lam) for one
    column, and with get_atomic_rfs ((I, Q, U, V, lam), rf) where rf is
    d I / d log gf [len(loggf_ids) + len(lam_ids), nlw] (the wavelength-shift
    rows are zero, as in the reference).  With get_populations the result is
    (that, populations): a tuple with one Populations per active atom (empty
    for an LTE run).  A run directory with an ACTIVE atom is solved in NLTE
    (rhf1d's Iterate and final ray) and returns the emergent spectrum of the
    converged atom.  Arrays own their memory and the inputs are not
    modified."""
    fudge = _fudge(fudge_wave, fudge_value)
    if loggf_ids is not None and loggf_values is not None and \
            len(loggf_ids) != len(loggf_values):
        raise ValueError("Different number of 'loggf_ids' and 'loggf_values'.")
    if lam_ids is not None and lam_values is not None and len(lam_ids) != len(lam_values):
        raise ValueError("Different number of 'lam_ids' and 'lam_values'.")
    if atomic_number is not None and atomic_abundance is not None and \
            len(atomic_number) != len(atomic_abundance):
        raise ValueError("Different number of 'atomic_number' and 'atomic_abundance'.")
    atm = np.ascontiguousarray(atmosphere, dtype=np.float64)
    if atm.ndim != 2 or atm.shape[0] < 9:
        raise ValueError("atmosphere must be [>=9, Ndep]")
    skey, s = _setup(cwd, atomic_number, atomic_abundance, loggf_ids, loggf_values,
                     lam_ids, lam_values)
    if any(a.active for a in s.atoms):
        if get_atomic_rfs:
            raise NotImplementedError("get_atomic_rfs with an ACTIVE atom is not implemented")
        e, nm = nlte_engine_for(cwd, wave, atomic_number, atomic_abundance, loggf_ids,
                                loggf_values, lam_ids, lam_values, fudge_wave=fudge_wave,
                                fudge_value=fudge_value)
        pops, _, _, st, nst = e.nlte_solve(nm, atm[None, :9], mu, atm_scale, stokes=True,
                                           nstar=True)
        lam = e.model.lam[e.model.out_index].copy()
        out = st[0, 0].copy(), st[0, 1].copy(), st[0, 2].copy(), st[0, 3].copy(), lam
        if get_populations:
            nd = atm.shape[1]
            return out, (Populations(nm.ID, int(nm.nlevel), nd, pops[0].copy(), nst[0].copy()),)
        return out
    e = engine_for(cwd, wave, atomic_number, atomic_abundance, loggf_ids, loggf_values,
                   lam_ids, lam_values, fudge_wave=fudge_wave, fudge_value=fudge_value)
    lam = e.model.lam[e.model.out_index].copy()
    if get_atomic_rfs:
        nlam = 0 if lam_ids is None else len(lam_ids)
        par = _model.atomic_rf_lines(e.model.setup, loggf_ids, nlam)
        if par.size == 0:
            raise ValueError("get_atomic_rfs needs loggf_ids or lam_ids")
        st, rf = e.rf_loggf(atm[None, :9], par, mu, atm_scale)
        out = st[0, 0].copy(), st[0, 1].copy(), st[0, 2].copy(), st[0, 3].copy(), lam
        out = out, np.ascontiguousarray(rf[0].T)
        return (out, ()) if get_populations else out
    st = e.synth_lte(atm[None, :9], mu, atm_scale)
    out = st[0, 0].copy(), st[0, 1].copy(), st[0, 2].copy(), st[0, 3].copy(), lam
    return (out, ()) if get_populations else out
