"""Host tables of the hydrostatic-equilibrium builder: pyrh's hse()
(pyrh.pyx:427-487, rh/rhf1d/pyrh_hse.c:67-399).

hse() reads the run directory as rhf1d() does, then treats every atom and
molecule as passive (rh/readatom.c:90, rh/readmolecule.c:85), reads no
Kurucz lines (pyrh_hse.c:137), puts hydrogen in LTE (:119) and evaluates
the background at the wavelengths SortLambda makes of {LAMBDA_REF, 500 nm}
(:197-200).  pyrh_Background (rh/rhf1d/pyrh_background.c:144-457) resets
its total opacity at every wavelength of that grid, so the opacity the
pressure integration uses is the one at the last (longest) of them.

build_hse returns the background model on that grid (uploaded with
rhg_open like any model) and the rhg_hse_t tables: every element of the
abundance table for the electron-density Newton step of get_ne
(pyrh_background.c:459-549, rh/solvene.c:145-211), hydrogen first, and the
Thomson and Rayleigh cross-sections and fudge factors at that wavelength.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, replace

import numpy as np

from . import constants as K
from .model import Model, RunSetup, build_model, linear
from .nlte import _rayleigh_atom, _rayleigh_h2
from .rhdata import RHInputError

HSE_LAMBDA = 500.0      # rh/rhf1d/pyrh_hse.c:197
NMAX_HSE_ITER = 50      # rh/rhf1d/pyrh_hse.c:63
HSE_ITER_LIMIT = 1e-2   # rh/rhf1d/pyrh_hse.c:269, 361


@dataclass
class HSEModel:
    model: Model
    nel: int
    el_nstage: np.ndarray
    el_pf_ptr: np.ndarray
    el_logpf: np.ndarray
    el_ionpot: np.ndarray
    el_abund: np.ndarray
    lam_index: int
    he_atom: int
    ray_H: float
    ray_He: float
    ray_H2: float
    sigma_thomson: float
    hmin_fudge: float
    scatt_fudge: float
    nmax_iter: int = NMAX_HSE_ITER
    iter_limit: float = HSE_ITER_LIMIT


def check_scale(atm_scale):
    """hse() integrates on the tau500 (0) and height (2) scales.  On the
    column-mass scale (1) the reference's pressure step is never set
    (pyrh_hse.c:282-290 and :312-328 cover only the other two), so its
    result is undefined; that scale is rejected."""
    if atm_scale == 1:
        raise NotImplementedError("hse on the column-mass scale: the reference leaves the "
                                  "pressure step undefined there (pyrh_hse.c:282-290)")
    if atm_scale not in (0, 2):
        raise ValueError(f"atm_scale must be 0 (tau500) or 2 (height), got {atm_scale}")


def _fudge_at(fudge_wave, fudge_value, lam):
    """H- and scattering fudge factors at lam, Linear over the fudge
    wavelengths (rh/rhf1d/pyrh_background.c:276-281, 387-392; rows 0 and 1
    of fudge_value, pyrh_hse.c:124-137)."""
    if fudge_wave is None and fudge_value is None:
        return 1.0, 1.0
    if fudge_wave is None or fudge_value is None:
        raise ValueError("fudge_wave and fudge_value go together")
    fw = np.ascontiguousarray(fudge_wave, np.float64).ravel()
    fv = np.ascontiguousarray(fudge_value, np.float64)
    if fv.ndim != 2 or fv.shape[0] != 3 or fv.shape[1] != fw.size or fw.size < 2:
        raise ValueError("fudge_value must be [3][len(fudge_wave)] with at least two wavelengths")
    return float(linear(fw, fv[0], lam)), float(linear(fw, fv[1], lam))


def build_hse(setup: RunSetup, fudge_wave=None, fudge_value=None) -> HSEModel:
    """Background model and rhg_hse_t tables of pyrh's hse() for the run
    directory behind `setup` (pyrh_amd.model.open_run, with the abundances
    hse() is given)."""
    kw = setup.kw
    if kw.bool("NON_ICE"):
        raise NotImplementedError("NON_ICE = TRUE")
    atoms = [replace(a, active=False) for a in setup.atoms]
    if not atoms or atoms[0].ID.strip() != "H":
        raise RHInputError("hydrogen must be the first atom of the atom list")
    hs = replace(setup, atoms=atoms, rlk=[])
    mdl = build_model(hs, np.array([HSE_LAMBDA]), check=False)
    lam_index = mdl.lam.size - 1
    lam = mdl.lam[lam_index:]

    el = setup.elements
    npf = el.Tpf.size
    elems = [e for e in range(len(el.abund)) if e == 0 or el.abund[e] != 0.0]
    nstage, ptr, logpf, ionp = [], [], [], []
    off = 0
    for e in elems:
        ns = int(el.nstage[e])
        nstage.append(ns)
        ptr.append(off)
        logpf.append(np.asarray(el.logpf[e], np.float64).ravel())
        ip = np.zeros(ns * npf)
        ip[:ns] = el.ionpot[e]
        ionp.append(ip)
        off += ns * npf

    vchar = kw.float("VMICRO_CHAR") * K.KM_TO_M
    he = [k for k, a in enumerate(atoms) if a.ID == "HE"]
    he_atom = he[0] if he else -1
    sigma_T = 8.0 * K.PI / 3.0 * math.pow(
        K.Q_ELECTRON / (math.sqrt(4.0 * K.PI * K.EPSILON_0) * (math.sqrt(K.M_ELECTRON) * K.CLIGHT)), 4)
    hf, sf = _fudge_at(fudge_wave, fudge_value, float(lam[0]))
    return HSEModel(
        model=mdl, nel=len(elems), el_nstage=np.array(nstage, np.int32),
        el_pf_ptr=np.array(ptr, np.int32), el_logpf=np.concatenate(logpf),
        el_ionpot=np.concatenate(ionp), el_abund=np.array([el.abund[e] for e in elems]),
        lam_index=int(lam_index), he_atom=he_atom,
        ray_H=float(_rayleigh_atom(atoms[0], lam, vchar)[0]),
        ray_He=float(_rayleigh_atom(atoms[he_atom], lam, vchar)[0]) if he_atom >= 0 else 0.0,
        ray_H2=float(_rayleigh_h2(lam)[0]), sigma_thomson=sigma_T,
        hmin_fudge=hf, scatt_fudge=sf)
