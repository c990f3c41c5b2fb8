"""ctypes binding of the gfx950 library (include/pyrh_amd.h).

There is no CPU fallback: if the library is missing or no GPU is present,
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
from pathlib import Path

import numpy as np

from .abi import ColumnsStruct, ModelStruct, RhgColumns, RhgModel

# the in-tree build; PYRH_AMD_LIB names another build of the same library
LIB_PATH = Path(os.environ.get("PYRH_AMD_LIB", Path(__file__).resolve().parent / "libpyrh_amd.so"))
_LIB = None

RHG_NSTAGE = 4
RHG_NWORK = 12
RF_NPARAM = 6
RF_PARAMS = ("T", "vlos", "B", "gamma", "chi", "vmic")
RF_ROWS = (1, 3, 5, 6, 7, 4)        # atmosphere rows of the parameters
STAGES = ("state", "geometry", "continuum", "formal")
# the four intervals rhg_last_timings reports after rhg_rf_lte
RF_STAGES = ("synthesis_inputs", "state_tangents", "forward", "reverse", "chain")
# rhg_nlte_timings slots (RHG_NLTE_NTIMING)
NLTE_TIMINGS = ("setup_ms", "scatter_bezier_ms", "iter_bezier_ms", "stateq_ms", "ng_ms",
                "total_ms", "rt_passes", "rt_pairs", "scatter_feautrier_ms", "iter_feautrier_ms",
                "cols_no_op", "cols_op")


# rhg_tune keys, in enum order
TUNE_KEYS = ("line_table", "column_sort", "wave_order")


class EngineError(RuntimeError):
    pass


def _share_hip_runtime():
    """PyTorch-ROCm bundles its own libamdhip64.so.7 with the same soname as
    the system one this library links to; the first one loaded serves the
    whole process, and torch cannot initialise on another build of it.
    Loading torch first (when installed) makes the library use torch's HIP
    runtime, so both work in one process and share device pointers."""
    if "torch" not in sys.modules:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass


def lib():
    """Load libpyrh_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        _share_hip_runtime()
        if not LIB_PATH.exists():
            raise EngineError(f"{LIB_PATH} is not built; run pyrh_amd.build.build_library() "
                              f"(or __graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        L.rhg_last_error.restype = C.c_char_p
        L.rhg_abi_version.restype = C.c_int32
        L.rhg_model_sizeof.restype = C.c_int32
        L.rhg_columns_sizeof.restype = C.c_int32
        L.rhg_open.argtypes = [C.POINTER(RhgModel), C.c_int, C.POINTER(C.c_void_p)]
        L.rhg_close.argtypes = [C.c_void_p]
        PD = C.POINTER(C.c_double)
        L.rhg_synth_lte.argtypes = [C.c_void_p, C.POINTER(RhgColumns), C.c_int32, C.c_int32, PD]
        L.rhg_synth_lte_noalloc.argtypes = [C.c_void_p, C.POINTER(RhgColumns), C.c_int32, PD]
        PI = C.POINTER(C.c_int32)
        L.rhg_set_grids.argtypes = [C.c_void_p, C.c_int32, PI, PI]
        L.rhg_synth_lte_grids.argtypes = [C.c_void_p, C.POINTER(RhgColumns), C.c_int32,
                                          C.c_int32, PI, PD]
        L.rhg_synth_lte_diag.argtypes = [C.c_void_p, C.POINTER(RhgColumns), C.c_int32, PD, PD, PD]
        L.rhg_last_timings.argtypes = [C.c_void_p, PD]
        L.rhg_set_tuning.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.rhg_get_tuning.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
        L.rhg_rf_timings.argtypes = [C.c_void_p, PD]
        L.rhg_count_work.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        L.rhg_rf_lte.argtypes = [C.c_void_p, C.POINTER(RhgColumns), C.c_int32, C.c_int32, PD,
                                 C.c_void_p]
        L.rhg_rf_lte_params.argtypes = [C.c_void_p, C.POINTER(RhgColumns), C.c_int32, C.c_int32,
                                        C.c_int32, C.POINTER(C.c_int32), PD, C.c_void_p]
        L.rhg_rf_loggf.argtypes = [C.c_void_p, C.POINTER(RhgColumns), C.c_int32,
                                   C.POINTER(C.c_int32), C.c_int32, PD, PD]
        L.rhg_nlte_solve.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(RhgColumns), C.c_int32,
                                     PD, C.POINTER(C.c_int32), PD, PD, PD]
        L.rhg_nlte_timings.argtypes = [C.c_void_p, PD]
        L.rhg_hse.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int32,
                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.rhg_hse_timing.argtypes = [C.c_void_p, PD, C.POINTER(C.c_int32)]
        L.rhg_ne_from_nh.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
        L.rhg_scales.argtypes = [C.c_void_p, C.POINTER(RhgColumns), C.c_int32, PD, PD, PD]
        if L.rhg_model_sizeof() != C.sizeof(RhgModel) or \
                L.rhg_columns_sizeof() != C.sizeof(RhgColumns):
            raise EngineError("rhg_model_t / rhg_columns_t layout mismatch between the "
                              "library and pyrh_amd.abi")
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise EngineError(f"pyrh_amd error {rc}: {lib().rhg_last_error().decode()}")


def _pd(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _pi(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int32))


def grid_table(positions):
    """(grid_ptr [ngrid + 1], grid_out) int32 of rhg_set_grids from a list of
    per-grid output positions."""
    parts = [np.asarray(p, dtype=np.int64).ravel() for p in positions]
    ptr = np.zeros(len(parts) + 1, np.int32)
    ptr[1:] = np.cumsum([p.size for p in parts])
    flat = np.concatenate(parts).astype(np.int32) if parts else np.zeros(0, np.int32)
    return ptr, np.ascontiguousarray(flat)


class Engine:
    """One device context holding one model (run directory + wavelength
    grid)."""

    def __init__(self, model, device: int = 0):
        self.model = model
        self._ms = ModelStruct(model)
        h = C.c_void_p()
        _check(lib().rhg_open(self._ms.ptr(), int(device), C.byref(h)))
        self._h = h
        self.nout = model.scalars["nspect"] - 1

    def close(self):
        if getattr(self, "_h", None):
            lib().rhg_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synth_lte(self, atmos, mu=1.0, atm_scale=0, ndep=None, precision=0, diag=False):
        """Emergent Stokes [ncol, 4, nout] float64 for host atmospheres
        [ncol, 9, ndep_max] in pyrh units."""
        cs = ColumnsStruct(atmos, ndep, atm_scale, mu)
        ncol, nd = cs.struct.ncol, cs.struct.ndep_max
        out = np.empty((ncol, 4, self.nout))
        if diag:
            h = np.zeros((ncol, nd))
            cr = np.zeros((ncol, nd))
            _check(lib().rhg_synth_lte_diag(self._h, cs.ptr(), int(precision), _pd(out),
                                            _pd(h), _pd(cr)))
            return out, h, cr
        _check(lib().rhg_synth_lte(self._h, cs.ptr(), int(precision), 0, _pd(out)))
        return out

    def synth_lte_device(self, d_atmos, d_ndep, ncol, ndep_max, d_out, mu=1.0, atm_scale=0,
                         precision=0):
        """Same, with device pointers (ints) for atmosphere, depth counts and
        output (e.g. from torch tensors); no host copies."""
        s = RhgColumns()
        s.ncol, s.ndep_max = int(ncol), int(ndep_max)
        s.ndep = C.cast(C.c_void_p(d_ndep), C.POINTER(C.c_int32))
        s.atmos = C.cast(C.c_void_p(d_atmos), C.POINTER(C.c_double))
        s.atm_scale, s.mu = int(atm_scale), float(mu)
        _check(lib().rhg_synth_lte(self._h, C.byref(s), int(precision), 1,
                                   C.cast(C.c_void_p(d_out), C.POINTER(C.c_double))))

    def set_grids(self, positions):
        """Wavelength grids of this context (rhg_set_grids): positions[g] are
        grid g's positions on the model's output wavelengths, increasing."""
        ptr, flat = grid_table(positions)
        _check(lib().rhg_set_grids(self._h, len(positions), _pi(ptr), _pi(flat)))
        self.grids = [np.asarray(p, np.int64).ravel().copy() for p in positions]

    def synth_lte_grids(self, atmos, col_grid, mu=1.0, atm_scale=0, ndep=None, precision=0):
        """synth_lte for columns on different grids of set_grids (column i on
        grid col_grid[i]), in one launch per stage: Stokes [ncol, 4, nout] on
        the model's output wavelengths, zero outside each column's grid."""
        cs = ColumnsStruct(atmos, ndep, atm_scale, mu)
        ncol = cs.struct.ncol
        g = np.ascontiguousarray(col_grid, dtype=np.int32)
        if g.shape != (ncol,):
            raise ValueError("col_grid must hold one grid id per column")
        out = np.empty((ncol, 4, self.nout))
        _check(lib().rhg_synth_lte_grids(self._h, cs.ptr(), int(precision), 0, _pi(g), _pd(out)))
        return out

    def synth_lte_grids_device(self, d_atmos, d_ndep, d_grid, ncol, ndep_max, d_out, mu=1.0,
                               atm_scale=0, precision=0):
        """Same, with device pointers (ints) for atmosphere, depth counts, grid
        ids and output; no host copies."""
        s = RhgColumns()
        s.ncol, s.ndep_max = int(ncol), int(ndep_max)
        s.ndep = C.cast(C.c_void_p(d_ndep), C.POINTER(C.c_int32))
        s.atmos = C.cast(C.c_void_p(d_atmos), C.POINTER(C.c_double))
        s.atm_scale, s.mu = int(atm_scale), float(mu)
        _check(lib().rhg_synth_lte_grids(self._h, C.byref(s), int(precision), 1,
                                         C.cast(C.c_void_p(d_grid), C.POINTER(C.c_int32)),
                                         C.cast(C.c_void_p(d_out), C.POINTER(C.c_double))))

    def rf_lte(self, atmos, mu=1.0, atm_scale=0, ndep=None, precision=0, params=None):
        """Stokes [ncol, 4, nout] and response functions
        [ncol, nparam, ndep_max, 4, nout] (d Stokes / d parameter at each
        depth; parameters T, v_LOS, B, gamma, chi, v_mic in the atmosphere's
        units, or the subset `params` of RF_PARAMS names in that order) for
        host atmospheres [ncol, 9, ndep_max]."""
        cs = ColumnsStruct(atmos, ndep, atm_scale, mu)
        ncol, nd = cs.struct.ncol, cs.struct.ndep_max
        out = np.empty((ncol, 4, self.nout))
        ids = np.arange(RF_NPARAM, dtype=np.int32) if params is None else \
            np.array([RF_PARAMS.index(p) for p in params], np.int32)
        rf = np.empty((ids.size, nd, 4, self.nout, ncol),
                      dtype=np.float32 if precision == 1 else np.float64)
        _check(lib().rhg_rf_lte_params(self._h, cs.ptr(), int(precision), 0, int(ids.size),
                                       ids.ctypes.data_as(C.POINTER(C.c_int32)), _pd(out),
                                       rf.ctypes.data_as(C.c_void_p)))
        return out, np.ascontiguousarray(np.moveaxis(rf, -1, 0))

    def rf_loggf(self, atmos, par_line, mu=1.0, atm_scale=0, ndep=None):
        """Stokes [ncol, 4, nout] and d I / d log gf [ncol, nout, npar] (the
        reference's analytic atomic response functions) for host atmospheres
        [ncol, 9, ndep_max]; par_line from pyrh_amd.model.atomic_rf_lines."""
        cs = ColumnsStruct(atmos, ndep, atm_scale, mu)
        ncol = cs.struct.ncol
        par = np.ascontiguousarray(par_line, dtype=np.int32)
        out = np.empty((ncol, 4, self.nout))
        rf = np.empty((ncol, self.nout, par.size))
        _check(lib().rhg_rf_loggf(self._h, cs.ptr(), int(par.size),
                                  par.ctypes.data_as(C.POINTER(C.c_int32)), 0, _pd(out), _pd(rf)))
        return out, rf

    def rf_lte_device(self, d_atmos, d_ndep, ncol, ndep_max, d_out, d_rf, mu=1.0, atm_scale=0,
                      precision=0, params=None):
        """rf_lte with device pointers: Stokes [ncol, 4, nout] float64 and the
        response functions in the library's layout [nparam, ndep_max, 4,
        nout, ncol] (float32 for precision 1, else float64) for all
        parameters or the subset `params` (RF_PARAMS names)."""
        s = RhgColumns()
        s.ncol, s.ndep_max = int(ncol), int(ndep_max)
        s.ndep = C.cast(C.c_void_p(d_ndep), C.POINTER(C.c_int32))
        s.atmos = C.cast(C.c_void_p(d_atmos), C.POINTER(C.c_double))
        s.atm_scale, s.mu = int(atm_scale), float(mu)
        ids = np.arange(RF_NPARAM, dtype=np.int32) if params is None else \
            np.array([RF_PARAMS.index(p) for p in params], np.int32)
        _check(lib().rhg_rf_lte_params(self._h, C.byref(s), int(precision), 1, int(ids.size),
                                       ids.ctypes.data_as(C.POINTER(C.c_int32)),
                                       C.cast(C.c_void_p(d_out), C.POINTER(C.c_double)),
                                       C.c_void_p(d_rf)))

    def nlte_solve(self, nm, atmos, mu=1.0, atm_scale=0, ndep=None, stokes=False, nstar=False):
        """NLTE populations of the active atom of nm (pyrh_amd.nlte.build_nlte;
        the engine must hold nm.lte) for host atmospheres [ncol, 9,
        ndep_max]: populations [ncol, nlevel, ndep_max], iterations [ncol]
        and the last maximum relative change [ncol]; with stokes=True also
        the emergent spectrum at mu afterwards [ncol, 4, nout], with
        nstar=True also the LTE populations (appended in that order)."""
        from .abi import NlteStruct
        ns = NlteStruct(nm)
        cs = ColumnsStruct(atmos, ndep, atm_scale, mu)
        ncol, nd = cs.struct.ncol, cs.struct.ndep_max
        pops = np.zeros((ncol, nm.nlevel, nd))
        niter = np.zeros(ncol, np.int32)
        dpops = np.zeros(ncol)
        st = np.zeros((ncol, 4, self.nout)) if stokes else None
        nst = np.zeros_like(pops) if nstar else None
        _check(lib().rhg_nlte_solve(self._h, C.addressof(ns.struct), cs.ptr(), 0, _pd(pops),
                                    niter.ctypes.data_as(C.POINTER(C.c_int32)), _pd(dpops),
                                    _pd(st), _pd(nst)))
        return (pops, niter, dpops) + ((st,) if stokes else ()) + ((nst,) if nstar else ())

    def nlte_solve_device(self, nm, d_atmos, d_ndep, ncol, ndep_max, d_pops, d_niter, d_dpops,
                          mu=1.0, atm_scale=0, d_stokes=None):
        """nlte_solve with device pointers (ints) for the atmospheres, depth
        counts and the outputs (d_stokes None: no emergent spectrum)."""
        from .abi import NlteStruct
        ns = NlteStruct(nm)
        s = RhgColumns()
        s.ncol, s.ndep_max = int(ncol), int(ndep_max)
        s.ndep = C.cast(C.c_void_p(d_ndep), C.POINTER(C.c_int32))
        s.atmos = C.cast(C.c_void_p(d_atmos), C.POINTER(C.c_double))
        s.atm_scale, s.mu = int(atm_scale), float(mu)
        _check(lib().rhg_nlte_solve(self._h, C.addressof(ns.struct), C.byref(s), 1,
                                    C.cast(C.c_void_p(d_pops), C.POINTER(C.c_double)),
                                    C.cast(C.c_void_p(d_niter), C.POINTER(C.c_int32)),
                                    C.cast(C.c_void_p(d_dpops), C.POINTER(C.c_double)),
                                    None if d_stokes is None else
                                    C.cast(C.c_void_p(d_stokes), C.POINTER(C.c_double)), None))

    def hse(self, hm, scale, temp, pg_top=0.1, atm_scale=0, ndep=None):
        """Hydrostatic equilibrium (pyrh's hse()) for the HSE model hm
        (pyrh_amd.hse.build_hse; the engine must hold hm.model) on host
        columns: scale, temp [ncol, ndep_max] (1-D: one column; scale may be
        one row for all), pg_top scalar or [ncol].  Returns ne, nHtot, rho,
        pg [ncol, ndep_max] in SI units and the pressure iterations per
        layer."""
        from .abi import HseStruct
        temp = np.ascontiguousarray(np.atleast_2d(temp), np.float64)
        ncol, nd = temp.shape
        scale = np.ascontiguousarray(np.broadcast_to(np.atleast_2d(np.asarray(scale, np.float64)),
                                                     temp.shape))
        ndep = (np.full(ncol, nd, np.int32) if ndep is None
                else np.ascontiguousarray(ndep, np.int32))
        if ndep.shape != (ncol,):
            raise ValueError("ndep must have one entry per column")
        pgt = np.ascontiguousarray(np.broadcast_to(np.asarray(pg_top, np.float64), (ncol,)))
        hs = HseStruct(hm)
        out = [np.zeros((ncol, nd)) for _ in range(4)]
        niter = np.zeros((ncol, nd), np.int32)
        _check(lib().rhg_hse(self._h, C.addressof(hs.struct), ncol, nd, ndep.ctypes.data,
                             int(atm_scale), scale.ctypes.data, temp.ctypes.data, pgt.ctypes.data, 0,
                             *(o.ctypes.data for o in out), niter.ctypes.data))
        return (*out, niter)

    def hse_device(self, hm, d_scale, d_temp, d_pg_top, d_ndep, ncol, ndep_max, d_ne, d_nHtot,
                   d_rho, d_pg, atm_scale=0, d_niter=None):
        """hse with device pointers (ints) for every array; d_niter None
        skips the iteration counts."""
        from .abi import HseStruct
        hs = HseStruct(hm)
        _check(lib().rhg_hse(self._h, C.addressof(hs.struct), int(ncol), int(ndep_max), d_ndep,
                             int(atm_scale), d_scale, d_temp, d_pg_top, 1, d_ne, d_nHtot, d_rho,
                             d_pg, d_niter))

    def ne_from_nh(self, hm, temp, nH, ndep=None):
        """pyrh's get_ne_from_nH on host columns: temp [K] and nH [cm^-3]
        [ncol, ndep_max] (1-D: one column).  Returns n_e [cm^-3] of the same
        shape (2-D)."""
        from .abi import HseStruct
        temp = np.ascontiguousarray(np.atleast_2d(temp), np.float64)
        nH = np.ascontiguousarray(np.atleast_2d(nH), np.float64)
        if nH.shape != temp.shape:
            raise ValueError("temp and nH must have the same shape")
        ncol, nd = temp.shape
        ndep = (np.full(ncol, nd, np.int32) if ndep is None
                else np.ascontiguousarray(ndep, np.int32))
        hs = HseStruct(hm)
        ne = np.zeros((ncol, nd))
        _check(lib().rhg_ne_from_nh(self._h, C.addressof(hs.struct), ncol, nd, ndep.ctypes.data,
                                    temp.ctypes.data, nH.ctypes.data, 0, ne.ctypes.data))
        return ne

    def scales(self, atmos, atm_scale=0, ndep=None):
        """pyrh's get_scales on host atmospheres [ncol, 9, ndep_max] (the
        engine's model holds the single reference wavelength): tau, height
        [m], cmass [kg m^-2], each [ncol, ndep_max]."""
        cs = ColumnsStruct(atmos, ndep, atm_scale, 1.0)
        ncol, nd = cs.struct.ncol, cs.struct.ndep_max
        out = [np.zeros((ncol, nd)) for _ in range(3)]
        _check(lib().rhg_scales(self._h, cs.ptr(), 0, *(_pd(o) for o in out)))
        return tuple(out)

    def hse_timing(self):
        """(kernel ms, lanes per column) of the last hse call."""
        ms = C.c_double()
        lanes = C.c_int32()
        _check(lib().rhg_hse_timing(self._h, C.byref(ms), C.byref(lanes)))
        return ms.value, lanes.value

    def nlte_timings(self):
        """Milliseconds and work counts of the last NLTE solve (keys of
        NLTE_TIMINGS)."""
        t = np.zeros(len(NLTE_TIMINGS))
        _check(lib().rhg_nlte_timings(self._h, _pd(t)))
        return dict(zip(NLTE_TIMINGS, t))

    def rf_timings(self):
        """Per-interval kernel milliseconds of the last rf_lte call."""
        t = np.zeros(len(RF_STAGES))
        _check(lib().rhg_rf_timings(self._h, _pd(t)))
        return dict(zip(RF_STAGES, t))

    def count_work(self):
        """Work census of the last synthesis (slots RHG_WORK_*, see
        pyrh_amd.roofline.WORK_SLOTS)."""
        cnt = np.zeros(RHG_NWORK, np.int64)
        _check(lib().rhg_count_work(self._h, cnt.ctypes.data_as(C.POINTER(C.c_int64))))
        return cnt

    def set_tuning(self, key, value):
        """Set one launch-configuration knob (TUNE_KEYS; 0 or 1)."""
        _check(lib().rhg_set_tuning(self._h, TUNE_KEYS.index(key), int(value)))

    def tuning(self):
        """The context's launch configuration {knob: value}."""
        out = {}
        for i, k in enumerate(TUNE_KEYS):
            v = C.c_int32()
            _check(lib().rhg_get_tuning(self._h, i, C.byref(v)))
            out[k] = int(v.value)
        return out

    def timings(self):
        """Per-stage kernel milliseconds of the last call."""
        t = np.zeros(RHG_NSTAGE)
        _check(lib().rhg_last_timings(self._h, _pd(t)))
        return dict(zip(STAGES, t))
