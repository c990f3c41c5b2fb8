"""ctypes mirror of the C ABI structs declared in include/pyrh_amd.h."""
from __future__ import annotations

import ctypes as C

import numpy as np

P_D = C.POINTER(C.c_double)
P_I = C.POINTER(C.c_int32)

# (name, ctype, kind) in declaration order; kind: 's' scalar, 'd'/'i' array
MODEL_FIELDS = [
    ("abi_version", C.c_int32, "s"),
    ("stokes_mode", C.c_int32, "s"), ("magneto_optical", C.c_int32, "s"),
    ("vmacro_tresh", C.c_double, "s"), ("vmicro_char", C.c_double, "s"),
    ("wght_per_H", C.c_double, "s"), ("totalAbund", C.c_double, "s"),
    ("gravity", C.c_double, "s"),
    ("nspect", C.c_int32, "s"), ("ref_index", C.c_int32, "s"),
    ("lamconst", P_D, "d"), ("solver", P_I, "i"),
    ("lam_line_ptr", P_I, "i"), ("lam_line_idx", P_I, "i"),
    ("lam_cont_ptr", P_I, "i"), ("lam_cont_idx", P_I, "i"), ("lam_cont_alpha", P_D, "d"),
    ("natom", C.c_int32, "s"),
    ("atom_level_ptr", P_I, "i"), ("level_E", P_D, "d"), ("level_g", P_D, "d"),
    ("level_stage", P_I, "i"), ("atom_abund", P_D, "d"),
    ("ncont", C.c_int32, "s"),
    ("cont_atom", P_I, "i"), ("cont_i", P_I, "i"), ("cont_j", P_I, "i"),
    ("nnuclei", C.c_int32, "s"),
    ("nucl_atom", P_I, "i"), ("nucl_elem", P_I, "i"), ("nucl_abund", P_D, "d"),
    ("nmol", C.c_int32, "s"),
    ("mol_fit", P_I, "i"), ("mol_charge", P_I, "i"), ("mol_nnuclei", P_I, "i"),
    ("mol_ediss", P_D, "d"), ("mol_tmin", P_D, "d"), ("mol_tmax", P_D, "d"),
    ("mol_neqc", P_I, "i"), ("mol_eqc", P_D, "d"), ("mol_nel", P_I, "i"),
    ("mol_el_nucl", P_I, "i"), ("mol_el_count", P_I, "i"),
    ("mol_h2", C.c_int32, "s"), ("mol_oh", C.c_int32, "s"), ("mol_ch", C.c_int32, "s"),
    ("npf", C.c_int32, "s"), ("Tpf", P_D, "d"),
    ("nel", C.c_int32, "s"),
    ("el_nstage", P_I, "i"), ("el_pf_ptr", P_I, "i"), ("el_logpf", P_D, "d"),
    ("el_ionpot", P_D, "d"), ("el_abund", P_D, "d"), ("el_weight", P_D, "d"),
    ("nline", C.c_int32, "s"),
    ("line_el", P_I, "i"), ("line_stage", P_I, "i"), ("line_lambda0", P_D, "d"),
    ("line_Ei", P_D, "d"), ("line_hc_la", P_D, "d"), ("line_Bijhc4pi", P_D, "d"),
    ("line_twohnu3_c2", P_D, "d"), ("line_Grad", P_D, "d"), ("line_GStark", P_D, "d"),
    ("line_GvdW", P_D, "d"), ("line_vdw", P_I, "i"), ("line_cross", P_D, "d"),
    ("line_alpha", P_D, "d"), ("line_larmor", P_D, "d"), ("line_polar", P_I, "i"),
    ("line_comp_ptr", P_I, "i"), ("comp_q", P_I, "i"), ("comp_shift", P_D, "d"),
    ("comp_strength", P_D, "d"),
    ("hminff_theta", P_D, "d"), ("hminff_kappa", P_D, "d"),
    ("h2mff_theta", P_D, "d"), ("h2mff_kappa", P_D, "d"),
    ("h2pff_temp", P_D, "d"), ("h2pff_kappa", P_D, "d"),
    ("oh_T", P_D, "d"), ("oh_cross", P_D, "d"),
    ("ch_T", P_D, "d"), ("ch_cross", P_D, "d"),
]

ABI_VERSION = 1


class RhgModel(C.Structure):
    _fields_ = [(n, t) for n, t, _ in MODEL_FIELDS]


class RhgColumns(C.Structure):
    _fields_ = [("ncol", C.c_int32), ("ndep_max", C.c_int32), ("ndep", P_I),
                ("atmos", P_D), ("atm_scale", C.c_int32), ("mu", C.c_double)]


class ModelStruct:
    """Owns contiguous copies of a Model's arrays and the struct pointing
    at them (the struct is valid while this object is alive)."""

    def __init__(self, model):
        self._keep = {}
        s = RhgModel()
        sc = dict(model.scalars)
        sc["abi_version"] = ABI_VERSION
        for name, ctype, kind in MODEL_FIELDS:
            if kind == "s":
                setattr(s, name, ctype(sc[name]).value)
            else:
                dt = np.float64 if kind == "d" else np.int32
                arr = np.ascontiguousarray(model.arrays[name], dtype=dt).ravel()
                self._keep[name] = arr
                setattr(s, name, arr.ctypes.data_as(ctype))
        self.struct = s

    def ptr(self):
        return C.byref(self.struct)


class ColumnsStruct:
    def __init__(self, atmos, ndep=None, atm_scale=0, mu=1.0):
        atmos = np.ascontiguousarray(atmos, dtype=np.float64)
        if atmos.ndim == 2:
            atmos = atmos[None]
        ncol, nrow, ndep_max = atmos.shape
        if nrow < 9:
            raise ValueError("atmosphere needs at least 9 rows")
        if nrow > 9:
            atmos = np.ascontiguousarray(atmos[:, :9])
        if ndep is None:
            ndep = np.full(ncol, ndep_max, np.int32)
        ndep = np.ascontiguousarray(ndep, dtype=np.int32)
        self.atmos, self.ndep = atmos, ndep
        s = RhgColumns()
        s.ncol, s.ndep_max = ncol, ndep_max
        s.ndep = ndep.ctypes.data_as(P_I)
        s.atmos = atmos.ctypes.data_as(P_D)
        s.atm_scale = int(atm_scale)
        s.mu = float(mu)
        self.struct = s

    def ptr(self):
        return C.byref(self.struct)


# rhg_nlte_t, in declaration order
NLTE_FIELDS = [
    ("nlevel", C.c_int32, "s"), ("atom_index", C.c_int32, "s"), ("he_atom", C.c_int32, "s"),
    ("E", P_D, "d"), ("g", P_D, "d"), ("stage", P_I, "i"), ("weight", C.c_double, "s"),
    ("ntrans", C.c_int32, "s"),
    ("tr_type", P_I, "i"), ("tr_i", P_I, "i"), ("tr_j", P_I, "i"),
    ("tr_nblue", P_I, "i"), ("tr_nlam", P_I, "i"), ("tr_ptr", P_I, "i"), ("tr_line", P_I, "i"),
    ("tr_wlam", P_D, "d"), ("tr_alpha", P_D, "d"),
    ("nline", C.c_int32, "s"),
    ("ln_lambda0", P_D, "d"), ("ln_Aji", P_D, "d"), ("ln_Bji", P_D, "d"), ("ln_Bij", P_D, "d"),
    ("ln_Grad", P_D, "d"), ("ln_vdw", P_I, "i"), ("ln_vdw_cross", P_D, "d"),
    ("ln_vdw_alpha", P_D, "d"), ("ln_vdw_cross_he", P_D, "d"),
    ("ln_stark", P_I, "i"), ("ln_stark_c", P_D, "d"), ("ln_stark_C", P_D, "d"),
    ("ln_stark_Cm", P_D, "d"),
    ("as_ptr", P_I, "i"), ("as_tr", P_I, "i"), ("as_bezier", P_I, "i"),
    ("ncoll", C.c_int32, "s"),
    ("coll_type", P_I, "i"), ("coll_i", P_I, "i"), ("coll_j", P_I, "i"), ("coll_ptr", P_I, "i"),
    ("coll_T", P_D, "d"), ("coll_c", P_D, "d"), ("coll_M", P_D, "d"),
    ("nrays", C.c_int32, "s"), ("muz", P_D, "d"), ("wmu", P_D, "d"),
    ("ray_H", P_D, "d"), ("ray_He", P_D, "d"), ("ray_H2", P_D, "d"),
    ("sigma_thomson", C.c_double, "s"),
    ("nmax_iter", C.c_int32, "s"), ("nmax_scatter", C.c_int32, "s"),
    ("ng_order", C.c_int32, "s"), ("ng_delay", C.c_int32, "s"), ("ng_period", C.c_int32, "s"),
    ("isum", C.c_int32, "s"), ("iter_limit", C.c_double, "s"),
]


class RhgNlte(C.Structure):
    _fields_ = [(n, t) for n, t, _ in NLTE_FIELDS]


class NlteStruct:
    """Owns contiguous copies of an NLTEModel's arrays and the rhg_nlte_t
    pointing at them."""

    def __init__(self, nm):
        self._keep = {}
        s = RhgNlte()
        counts = {"ntrans": int(nm.tr_type.size), "nline": int(nm.ln_lambda0.size),
                  "ncoll": int(nm.coll_type.size), "nrays": int(nm.muz.size)}
        for name, ctype, kind in NLTE_FIELDS:
            if kind == "s":
                v = counts[name] if name in counts else getattr(nm, name)
                setattr(s, name, ctype(v).value)
            else:
                dt = np.float64 if kind == "d" else np.int32
                arr = np.ascontiguousarray(getattr(nm, name), dtype=dt).ravel()
                if arr.size == 0:
                    arr = np.zeros(1, dtype=dt)
                self._keep[name] = arr
                setattr(s, name, arr.ctypes.data_as(ctype))
        self.struct = s

    def ptr(self):
        return C.byref(self.struct)


# rhg_hse_t, in declaration order
HSE_FIELDS = [
    ("nel", C.c_int32, "s"), ("el_nstage", P_I, "i"), ("el_pf_ptr", P_I, "i"),
    ("el_logpf", P_D, "d"), ("el_ionpot", P_D, "d"), ("el_abund", P_D, "d"),
    ("lam_index", C.c_int32, "s"), ("he_atom", C.c_int32, "s"),
    ("ray_H", C.c_double, "s"), ("ray_He", C.c_double, "s"), ("ray_H2", C.c_double, "s"),
    ("sigma_thomson", C.c_double, "s"),
    ("hmin_fudge", C.c_double, "s"), ("scatt_fudge", C.c_double, "s"),
    ("nmax_iter", C.c_int32, "s"), ("iter_limit", C.c_double, "s"),
]


class RhgHse(C.Structure):
    _fields_ = [(n, t) for n, t, _ in HSE_FIELDS]


class HseStruct:
    """Owns contiguous copies of an HSEModel's tables and the rhg_hse_t
    pointing at them."""

    def __init__(self, hm):
        self._keep = {}
        s = RhgHse()
        for name, ctype, kind in HSE_FIELDS:
            if kind == "s":
                setattr(s, name, ctype(getattr(hm, name)).value)
            else:
                dt = np.float64 if kind == "d" else np.int32
                arr = np.ascontiguousarray(getattr(hm, name), dtype=dt).ravel()
                if arr.size == 0:
                    arr = np.zeros(1, dtype=dt)
                self._keep[name] = arr
                setattr(s, name, arr.ctypes.data_as(ctype))
        self.struct = s

    def ptr(self):
        return C.byref(self.struct)
