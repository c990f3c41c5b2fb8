"""pyrh-compatible Python API on top of the GPU engine.

``compute1d`` keeps the signature and return tuple of pyrh.compute1d
(pyrh.pyx:537-668) for the LTE path; ``synth_batch`` is the batched form
that the GPU is built for.  Parsed inputs, models and device contexts are
cached per run directory and wavelength grid, so repeated calls do not
re-read files (the reference re-reads everything on every call,
rh/rhf1d/pyrh_compute1dray.c:130-301).
"""
from __future__ import annotations

import hashlib
import os

import numpy as np

from . import model as _model
from .engine import Engine

_SETUPS = {}
_ENGINES = {}


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


def engine_for(cwd, wave, atomic_number=None, atomic_abundance=None, loggf_ids=None,
               loggf_values=None, lam_ids=None, lam_values=None, device=0):
    skey, s = _setup(cwd, atomic_number, atomic_abundance, loggf_ids, loggf_values,
                     lam_ids, lam_values)
    key = (skey, _key(np.asarray(wave, np.float64)), device)
    e = _ENGINES.get(key)
    if e is None:
        m = _model.build_model(s, np.asarray(wave, np.float64))
        e = Engine(m, device)
        _ENGINES[key] = e
    return e


def synth_batch(cwd, mu, atm_scale, atmospheres, wave, ndep=None, precision="fp64",
                atomic_number=None, atomic_abundance=None, device=0):
    """Stokes I, Q, U, V [ncol, nlw] and the wavelengths [nlw] for a batch
    of columns [ncol, >=9, ndep_max] (pyrh row order and units)."""
    e = engine_for(cwd, wave, atomic_number, atomic_abundance, device=device)
    prec = {"fp64": 0, "fp32": 1}[precision]
    st = e.synth_lte(np.asarray(atmospheres, np.float64), mu, atm_scale, ndep, prec)
    lam = e.model.lam[e.model.out_index]
    return st[:, 0], st[:, 1], st[:, 2], st[:, 3], lam


def compute1d(cwd, mu, atm_scale, atmosphere, wave, loggf_ids=None, loggf_values=None,
              lam_ids=None, lam_values=None, fudge_wave=None, fudge_value=None,
              atomic_number=None, atomic_abundance=None, get_atomic_rfs=False,
              get_populations=False):
    """pyrh.compute1d (pyrh.pyx:537-668): returns (I, Q, U, V, lam) for one
    column, and with get_atomic_rfs ((I, Q, U, V, lam), rf) where rf is
    d I / d log gf [len(loggf_ids) + len(lam_ids), nlw] (the wavelength-shift
    rows are zero, as in the reference).  Arrays own their memory and the
    inputs are not modified."""
    if get_populations:
        raise NotImplementedError("get_populations applies to active (NLTE) atoms, which "
                                  "this LTE engine does not solve")
    if fudge_wave is not None or fudge_value is not None:
        raise NotImplementedError("opacity fudge factors are not implemented")
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
    e = engine_for(cwd, wave, atomic_number, atomic_abundance, loggf_ids, loggf_values,
                   lam_ids, lam_values)
    lam = e.model.lam[e.model.out_index].copy()
    if get_atomic_rfs:
        nlam = 0 if lam_ids is None else len(lam_ids)
        par = _model.atomic_rf_lines(e.model.setup, loggf_ids, nlam)
        if par.size == 0:
            raise ValueError("get_atomic_rfs needs loggf_ids or lam_ids")
        st, rf = e.rf_loggf(atm[None, :9], par, mu, atm_scale)
        out = st[0, 0].copy(), st[0, 1].copy(), st[0, 2].copy(), st[0, 3].copy(), lam
        return out, np.ascontiguousarray(rf[0].T)
    st = e.synth_lte(atm[None, :9], mu, atm_scale)
    return st[0, 0].copy(), st[0, 1].copy(), st[0, 2].copy(), st[0, 3].copy(), lam
