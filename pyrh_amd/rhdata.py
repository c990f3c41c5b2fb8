"""Readers for RH/pyrh input files: keyword.input, atoms.input,
molecules.input, kurucz.input and the files they name.

The semantics (defaults, comment handling, column offsets, unit
conversions, which quantities are derived at read time) follow the
reference readers cited per function.  Everything here runs once per
batch on the host; the per-column work happens on the GPU.
"""
from __future__ import annotations

import math
import os
import struct
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import constants as K
from .cscan import data_lines, sscanf


class RHInputError(RuntimeError):
    """Raised where the reference would print an error and exit()."""


# Atomic IDs and weights of the first 99 elements; the same table as
# rh/atomweights.h:20-46 (standard atomic weights), index = Z-1.
ELEMENTS = [
    ('H ', 1.008), ('HE', 4.003), ('LI', 6.939), ('BE', 9.013), ('B ', 10.810), ('C ', 12.010),
    ('N ', 14.010), ('O ', 16.000), ('F ', 19.000), ('NE', 20.180), ('NA', 22.990), ('MG', 24.310),
    ('AL', 26.980), ('SI', 28.090), ('P ', 30.980), ('S ', 32.070), ('CL', 35.450), ('AR', 39.950),
    ('K ', 39.100), ('CA', 40.080), ('SC', 44.960), ('TI', 47.900), ('V ', 50.940), ('CR', 52.000),
    ('MN', 54.940), ('FE', 55.850), ('CO', 58.940), ('NI', 58.710), ('CU', 63.550), ('ZN', 65.370),
    ('GA', 69.720), ('GE', 72.600), ('AS', 74.920), ('SE', 78.960), ('BR', 79.910), ('KR', 83.800),
    ('RB', 85.480), ('SR', 87.630), ('Y ', 88.910), ('ZR', 91.220), ('NB', 92.910), ('MO', 95.950),
    ('TC', 99.000), ('RU', 101.100), ('RH', 102.900), ('PD', 106.400), ('AG', 107.900), ('CD', 112.400),
    ('IN', 114.800), ('SN', 118.700), ('SB', 121.800), ('TE', 127.600), ('I ', 126.900), ('XE', 131.300),
    ('CS', 132.900), ('BA', 137.400), ('LA', 138.900), ('CE', 140.100), ('PR', 140.900), ('ND', 144.300),
    ('PM', 147.000), ('SM', 150.400), ('EU', 152.000), ('GD', 157.300), ('TB', 158.900), ('DY', 162.500),
    ('HO', 164.900), ('ER', 167.300), ('TM', 168.900), ('YB', 173.000), ('LU', 175.000), ('HF', 178.500),
    ('TA', 181.000), ('W ', 183.900), ('RE', 186.300), ('OS', 190.200), ('IR', 192.200), ('PT', 195.100),
    ('AU', 197.000), ('HG', 200.600), ('TL', 204.400), ('PB', 207.200), ('BI', 209.000), ('PO', 210.000),
    ('AT', 211.000), ('RN', 222.000), ('FR', 223.000), ('RA', 226.100), ('AC', 227.100), ('TH', 232.000),
    ('PA', 231.000), ('U ', 238.000), ('NP', 237.000), ('PU', 244.000), ('AM', 243.000), ('CM', 247.000),
    ('BK', 247.000), ('CF', 251.000), ('ES', 254.000),
]


def _find_element(elem_id: str) -> int:
    """First element whose 2-char ID occurs in ``elem_id`` (the
    ``strstr(id, elements[n].ID)`` lookups of rh/readatom.c:127 and
    rh/readmolecule.c:143); the abundance reader uses the reverse test."""
    for n, (eid, _) in enumerate(ELEMENTS):
        if eid in elem_id:
            return n
    return -1


# ---------------------------------------------------------------------------
# keyword.input
# ---------------------------------------------------------------------------

# (keyword, default, type) for the keywords that influence this path;
# rh/readinput.c:51-210.  REQUIRED keywords have no default.
_KEYWORDS = {
    "ABUND_FILE": ("", "optional"), "NRAYS": ("0", "optional"),
    "ATMOS_ITOP": ("none", "optional"), "ATOMS_FILE": ("", "required"),
    "BARKLEM_SP_DATA": ("", "optional"), "BARKLEM_PD_DATA": ("", "optional"),
    "BARKLEM_DF_DATA": ("", "optional"), "LS_LANDE": ("TRUE", "default"),
    "MOLECULES_FILE": ("", "required"), "NON_ICE": ("FALSE", "optional"),
    "N_MAX_SCATTER": ("5", "optional"), "I_SUM": ("0", "required"),
    "N_MAX_ITER": ("", "required"), "ITER_LIMIT": ("", "required"),
    "NG_DELAY": ("0", "optional"), "NG_ORDER": ("0", "optional"),
    "NG_PERIOD": ("1", "optional"), "PRD_N_MAX_ITER": ("3", "optional"),
    "PRD_ITER_LIMIT": ("1.0E-2", "optional"), "J_FILE": ("", "required"),
    "BACKGROUND_FILE": ("", "required"), "STARTING_J": ("", "required"),
    "HYDROGEN_LTE": ("FALSE", "default"), "HYDROSTATIC": ("FALSE", "default"),
    "KURUCZ_DATA": ("none", "optional"), "RLK_SCATTER": ("FALSE", "default"),
    "KURUCZ_PF_DATA": ("", "optional"), "SOLVE_NE": ("NONE", "default"),
    "OPACITY_FUDGE": ("none", "optional"), "DO_FUDGE": ("FALSE", "optional"),
    "METALLICITY": ("0.0", "default"), "VMICRO_CHAR": ("", "required"),
    "VMACRO_TRESH": ("0.1", "optional"), "LAMBDA_REF": ("500.0", "default"),
    "VACUUM_TO_AIR": ("0", "optional"), "STOKES_INPUT": ("none", "optional"),
    "B_STRENGTH_CHAR": ("0.0", "default"), "STOKES_MODE": ("NO_STOKES", "optional"),
    "MAGNETO_OPTICAL": ("TRUE", "default"),
    "BACKGROUND_POLARIZATION": ("FALSE", "default"),
    "S_INTERPOLATION": ("S_BEZIER3", "default"),
    "S_INTERPOLATION_STOKES": ("DELO_BEZIER3", "default"),
    "LIMIT_MEMORY": ("FALSE", "default"), "ALLOW_PASSIVE_BB": ("TRUE", "default"),
    "PRD_ANGLE_DEP": ("FALSE", "default"), "XRD": ("FALSE", "default"),
    "EDDINGTON": ("FALSE", "optional"), "ANGLE_SET": ("NO_SET", "optional"),
}
# Keywords the reference accepts but which only touch file output,
# threading or other subsystems; they are read and ignored here.
_IGNORED = {
    "BACKGROUND_RAY_FILE", "OLD_BACKGROUND", "PRD_NG_DELAY", "PRD_NG_ORDER",
    "PRD_NG_PERIOD", "NG_MOLECULES", "ATMOS_OUTPUT", "GEOMETRY_OUTPUT",
    "SPECTRUM_OUTPUT", "OPACITY_OUTPUT", "RADRATE_OUTPUT", "COLLRATE_OUTPUT",
    "DAMPING_OUTPUT", "COOLING_OUTPUT", "XDR_ENDIAN", "INTERPOLATE_3D",
    "PRINT_CPU", "N_THREADS",
}


@dataclass
class Keywords:
    values: dict

    def str(self, k):
        return self.values[k]

    def int(self, k):
        return sscanf(self.values[k], "%d")[1][0] if self.values[k] else 0

    def float(self, k):
        return sscanf(self.values[k], "%lf")[1][0]

    def bool(self, k):
        # setboolValue: TRUE iff the value contains "TRUE" (rh/readvalue.c:187)
        return "TRUE" in self.values[k]


def read_keywords(path) -> Keywords:
    """rh/readvalue.c:70-125 (``KEYWORD = value`` per non-comment line)."""
    if not os.path.exists(path):
        raise RHInputError(f"Unable to open inputfile {path}")
    vals = {k: d for k, (d, _) in _KEYWORDS.items()}
    seen = set()
    for line in data_lines(path):
        n, v = sscanf(line, "%s = %s")
        if n != 2:
            raise RHInputError(f"Missing input value for keyword in line {line.strip()!r}")
        kw, value = v
        if kw in _KEYWORDS:
            vals[kw] = value
            seen.add(kw)
        elif kw not in _IGNORED:
            pass  # the reference only warns about unknown keywords
    for k, (_, typ) in _KEYWORDS.items():
        if typ == "required" and k not in seen:
            raise RHInputError(f"Missing input value for required keyword {k}")
    return Keywords(vals)


# ---------------------------------------------------------------------------
# abundance.input and the Kurucz partition-function table (XDR)
# ---------------------------------------------------------------------------

@dataclass
class ElementTable:
    abund: np.ndarray          # [Nelem] number abundance relative to H
    abundance_set: np.ndarray  # [Nelem] bool
    weight: np.ndarray         # [Nelem] amu
    nstage: np.ndarray         # [Nelem]
    Tpf: np.ndarray            # [Npf]
    logpf: list                # [Nelem] -> [Nstage, Npf] natural log of pf
    ionpot: list               # [Nelem] -> [Nstage] in J
    totalAbund: float
    wght_per_H: float
    avgMolWght: float


def read_abundance(abund_path, pf_path, metallicity=0.0, Nabun=0,
                   atomic_id=None, atomic_abundance=None) -> ElementTable:
    """rh/abundance.c:72-226: abundances (dex if H = 12), metallicity
    scaling, and the XDR partition-function table (log pf, ionisation
    potentials converted from cm^-1 to J)."""
    nel = len(ELEMENTS)
    abund = np.zeros(nel)
    aset = np.zeros(nel, bool)
    weight = np.array([w for _, w in ELEMENTS])
    if not os.path.exists(abund_path):
        raise RHInputError(f"Unable to open input file {abund_path}")
    dex = False
    ida_reading = 1
    for line in data_lines(abund_path):
        n, v = sscanf(line, "%s %lf")
        if n != 2:
            raise RHInputError(f"Unable to read input file {abund_path}")
        eid, ab = v[0].upper(), v[1]
        if len(eid) == 1:
            eid += " "
        if Nabun:
            for ida in range(Nabun):
                if int(atomic_id[ida]) == ida_reading:
                    ab = float(atomic_abundance[ida])
                    break
        ida_reading += 1
        for m, (mid, _) in enumerate(ELEMENTS):
            if eid in mid:
                abund[m] = ab
                if "H " in eid and ab == 12.0:
                    dex = True
                aset[m] = True
                break
    metal = K.pow10(metallicity)

    raw = Path(pf_path).read_bytes()
    pos = 0

    def xi():
        nonlocal pos
        (v,) = struct.unpack_from(">i", raw, pos)
        pos += 4
        return v

    def xd(n):
        nonlocal pos
        a = np.frombuffer(raw, ">f8", n, pos).astype(np.float64)
        pos += 8 * n
        return a

    npf = xi()
    Tpf = xd(npf)
    nstage = np.zeros(nel, np.int32)
    logpf, ionpot = [], []
    total = avgw = 0.0
    for m in range(nel):
        if aset[m]:
            if dex:
                abund[m] = K.pow10(abund[m] - 12.0)
            if metal != 1.0 and ELEMENTS[m][0] != "H ":
                abund[m] *= metal
            total += abund[m]
            avgw += abund[m] * weight[m]
        xi()  # periodic-table index
        ns = xi()
        nstage[m] = ns
        pf = xd(ns * npf).reshape(ns, npf)
        ip = xd(ns)
        if aset[m]:
            # 'ionpot[i] *= (HPLANCK * CLIGHT) / CM_TO_M' (constant first)
            ip = ip * ((K.HPLANCK * K.CLIGHT) / K.CM_TO_M)
            # glibc log, element by element (numpy's vectorised log can
            # differ from the C library's in the last bit)
            pf = np.array([[math.log(x) for x in row] for row in pf])
        logpf.append(pf)
        ionpot.append(ip)
    return ElementTable(abund, aset, weight, nstage, Tpf, logpf, ionpot,
                        total, avgw, avgw / total)


# ---------------------------------------------------------------------------
# Atom models (rh/readatom.c:62-615)
# ---------------------------------------------------------------------------

@dataclass
class AtomLine:
    i: int
    j: int
    f: float
    shape: str
    Nlambda: int
    symmetric: bool
    qcore: float
    qwing: float
    vdw: str
    cvdWaals: tuple
    Grad: float
    cStark: float
    g_Lande_eff: float
    lambda0: float      # nm (vacuum)
    Aji: float
    Bji: float
    Bij: float
    Ncomponent: int = 1
    c_shift: tuple = (0.0,)
    c_fraction: tuple = (1.0,)


@dataclass
class AtomCont:
    i: int
    j: int
    alpha0: float
    Nlambda: int
    hydrogenic: bool
    lambda0: float        # edge, nm
    lam: np.ndarray       # grid, nm (ascending)
    alpha: np.ndarray     # explicit cross-sections [m^2] (empty if hydrogenic)


@dataclass
class Atom:
    ID: str
    path: str
    active: bool
    initial_solution: str
    periodic_table: int
    abundance: float
    weight: float
    E: np.ndarray      # J
    g: np.ndarray
    stage: np.ndarray
    label: list
    abo_level: list
    lines: list
    conts: list
    raw_lines: list = field(default_factory=list)  # remaining text (collisions)
    Nfixed: int = 0


def _atom_lines_after(path):
    return list(data_lines(path))


def read_atom(path, active, elements: ElementTable, initial="LTE_POPULATIONS") -> Atom:
    """Parse an RH .atom file the way readAtom does (rh/readatom.c:62-615)."""
    if not os.path.exists(path):
        raise RHInputError(f"Unable to open input file {path}")
    L = _atom_lines_after(path)
    p = 0

    def nxt():
        nonlocal p
        if p >= len(L):
            raise RHInputError("Reached end of input file before all data was read")
        s = L[p]
        p += 1
        return s

    n, v = sscanf(nxt(), "%2s")
    aid = v[0].upper()
    if len(aid) == 1:
        aid += " "
    pt = -1
    for m, (eid, _) in enumerate(ELEMENTS):
        if aid in eid:
            if elements.abundance_set[m]:
                pt = m
            break
    if pt < 0:
        raise RHInputError(f"No matching element in periodic table for element {aid} "
                           f"in file {path}, or abundance not specified")
    n, v = sscanf(nxt(), "%d %d %d %d")
    if n < 4:
        raise RHInputError(f"Unable to read input file {path} (checkpoint 2)")
    Nlevel, Nline, Ncont, Nfixed = v
    E = np.zeros(Nlevel)
    g = np.zeros(Nlevel)
    stage = np.zeros(Nlevel, np.int32)
    labels, abo = [], []
    for i in range(Nlevel):
        s = nxt()
        n, v = sscanf(s, "%lf %lf '%20c' %d %d %d")
        if n < 4:
            raise RHInputError(f"Unable to read input file {path} (checkpoint 3)")
        E[i] = v[0] * ((K.HPLANCK * K.CLIGHT) / K.CM_TO_M)   # 'E *= hc/CM_TO_M'
        g[i] = v[1]
        labels.append(v[2])
        stage[i] = v[3]
        abo.append(v[5] if n >= 6 else -1)
    if stage[Nlevel - 1] != stage[Nlevel - 2] + 1:
        raise RHInputError(f"Atomic model {aid} in file {path} does not have overlying continuum")

    C = 2 * K.PI * (K.Q_ELECTRON / K.EPSILON_0) * (K.Q_ELECTRON / K.M_ELECTRON) / K.CLIGHT
    lines = []
    for kr in range(Nline):
        n, v = sscanf(nxt(), "%d %d %lf %s %d %s %lf %lf %s %lf %lf %lf %lf %lf %lf %lf")
        if n < 15:
            raise RHInputError(f"Unable to read input file {path} (checkpoint 4)")
        j0, i0, f, shape, nlam, symm, qcore, qwing, vdw = v[:9]
        cvdw = list(v[9:13])
        Grad, cStark = v[13], v[14]
        gL = v[15] if n >= 16 else 0.0
        j, i = max(i0, j0), min(i0, j0)
        lam0 = (K.HPLANCK * K.CLIGHT) / (E[j] - E[i])
        Aji = C / (lam0 * lam0) * (g[i] / g[j]) * f
        Bji = (lam0 * lam0 * lam0) / (2.0 * K.HPLANCK * K.CLIGHT) * Aji
        Bij = (g[j] / g[i]) * Bji
        if not any(s in shape for s in ("PRD", "VOIGT", "GAUSS", "COMPOSIT")):
            raise RHInputError(f"Invalid value for line-shape string: {shape}")
        ln = AtomLine(i, j, f, shape, nlam, "ASYMM" not in symm, qcore, qwing, vdw,
                      tuple(cvdw), Grad, cStark, gL, lam0 / K.NM_TO_M, Aji, Bji, Bij)
        if "COMPOSIT" in shape:
            nc = sscanf(nxt(), "%d")[1][0]
            sh, fr = [], []
            for _ in range(nc):
                _, cv = sscanf(nxt(), "%lf %lf")
                sh.append(cv[0])
                fr.append(cv[1])
            ln.Ncomponent, ln.c_shift, ln.c_fraction = nc, tuple(sh), tuple(fr)
        lines.append(ln)

    conts = []
    for kr in range(Ncont):
        n, v = sscanf(nxt(), "%d %d %lf %d %s %lf")
        if n < 6:
            raise RHInputError(f"Unable to read input file {path} (checkpoint 5)")
        j0, i0, alpha0, nlam, nudep, lammin = v
        j, i = max(i0, j0), min(i0, j0)
        lam0 = (K.HPLANCK * K.CLIGHT) / (E[j] - E[i]) / K.NM_TO_M
        if "EXPLICIT" in nudep:
            lam = np.zeros(nlam)
            alp = np.zeros(nlam)
            for la in range(nlam - 1, -1, -1):
                _, cv = sscanf(nxt(), "%lf %lf")
                lam[la], alp[la] = cv
            if np.any(np.diff(lam) < 0):
                raise RHInputError(f"Wavelength for continuum {i} - {j} is not monotonous")
            conts.append(AtomCont(i, j, alpha0, nlam, False, lam0, lam, alp))
        elif "HYDROGENIC" in nudep:
            if lammin >= lam0:
                raise RHInputError(f"Minimum wavelength for continuum {j} -> {i} is larger "
                                   f"than continuum edge")
            # getLambdaCont (rh/getlambda.c:167-180): running sum
            lam = np.zeros(nlam)
            dl = (lam0 - lammin) / (nlam - 1)
            lam[0] = lammin
            for la in range(1, nlam):
                lam[la] = lam[la - 1] + dl
            conts.append(AtomCont(i, j, alpha0, nlam, True, lam0, lam, np.zeros(0)))
        else:
            raise RHInputError(f"Invalid value for wavelength dependence string: {nudep}")
    rest = L[p:]
    return Atom(aid, str(path), active, initial, pt, float(elements.abund[pt]),
                float(elements.weight[pt]), E, g, stage, labels, abo, lines, conts, rest,
                Nfixed=int(Nfixed))


def read_atoms_input(path, pyrh_path, elements) -> list:
    """atoms.input -> list of Atom (rh/readatom.c:800-921); model files are
    looked up in $PYRH_PATH/rh/Atoms/ as pyrh does."""
    if not os.path.exists(path):
        raise RHInputError(f"Unable to open input file {path}")
    L = list(data_lines(path))
    natom = sscanf(L[0], "%d")[1][0]
    atoms = []
    for n in range(natom):
        cnt, v = sscanf(L[1 + n], "%s %s %s %s")
        if cnt < 3:
            raise RHInputError(f"Unable to read input file {path}")
        fname, action, popskey = v[0], v[1], v[2]
        full = os.path.join(pyrh_path, "rh", "Atoms", fname)
        init = ("OLD_POPULATIONS" if "OLD_POPULATIONS" in popskey else
                "LTE_POPULATIONS" if "LTE_POPULATIONS" in popskey else
                "ZERO_RADIATION" if "ZERO_RADIATION" in popskey else None)
        if init is None:
            raise RHInputError(f"Unknown initial solution specified for atom {fname}")
        at = read_atom(full, "ACTIVE" in action, elements, init)
        if n == 0 and at.ID != "H ":
            raise RHInputError(f"First atomic model is not hydrogen: {at.ID}")
        for prev in atoms:
            if prev.ID in at.ID:
                raise RHInputError(f"Aready read atomic model for element {at.ID}")
        atoms.append(at)
    return atoms


# ---------------------------------------------------------------------------
# Molecules (rh/readmolecule.c:66-300, 930-1080)
# ---------------------------------------------------------------------------

FIT_TYPES = {"KURUCZ_70": 0, "KURUCZ_85": 1, "SAUVAL_TATUM_84": 2, "IRWIN_81": 3,
             "TSUJI_73": 4}


@dataclass
class Molecule:
    ID: str
    charge: int
    pt_index: list
    pt_count: list
    Nnuclei: int
    Ediss: float        # J
    fit: int
    Tmin: float
    Tmax: float
    pf_coef: list       # highest power first (as stored by the reference)
    eqc_coef: list
    line_files: list
    weight: float


def read_molecule(path, elements: ElementTable) -> Molecule:
    if not os.path.exists(path):
        raise RHInputError(f"Unable to open inputfile {path}")
    L = list(data_lines(path))
    mid = sscanf(L[0], "%s")[1][0]
    charge = sscanf(L[1], "%d")[1][0]
    if charge < 0 or charge > 1:
        raise RHInputError(f"Only neutral or singly charged positive molecules allowed: {mid}")
    toks = [t for t in L[2].replace(",", " ").split()]
    pt_index, pt_count = [], []
    for t in toks:
        cnt, v = sscanf(t, "%d%s")
        if cnt == 0:
            c, eid = 1, t
        else:
            c, eid = v[0], v[1]
        eid = eid.upper()
        if len(eid) == 1:
            eid += " "
        m = _find_element(eid)
        if m < 0 or not elements.abundance_set[m]:
            raise RHInputError(f"Could not find element {eid} of molecule {mid}")
        pt_index.append(m)
        pt_count.append(c)
    Ediss = sscanf(L[3], "%lf")[1][0] * K.EV
    fitstr = sscanf(L[4], "%s")[1][0]
    fit = None
    for name, code in FIT_TYPES.items():
        if name in fitstr:
            fit = code
            break
    if fit is None:
        raise RHInputError(f"Unable to regognize keyword {fitstr} in file {path}")
    Tmin, Tmax = sscanf(L[5], "%lf %lf")[1]

    def coefs(line):
        t = line.split()
        npf = int(t[0])
        c = [float(x) for x in t[1:1 + npf]]
        return c[::-1]  # the reference stores them reversed (index n = Npf-1 .. 0)

    pf = coefs(L[6])
    eqc = coefs(L[7])
    files = [sscanf(s, "%s")[1][0] for s in L[8:]]
    weight = 0.0
    for ix, c in zip(pt_index, pt_count):
        weight += c * elements.weight[ix]
    return Molecule(mid, charge, pt_index, pt_count, sum(pt_count), Ediss, fit, Tmin, Tmax,
                    pf, eqc, files, weight)


def read_molecules_input(path, pyrh_path, elements) -> list:
    if not os.path.exists(path):
        raise RHInputError(f"Unable to open input file {path}")
    L = list(data_lines(path))
    nmol = sscanf(L[0], "%d")[1][0]
    mols = []
    for n in range(nmol):
        cnt, v = sscanf(L[1 + n], "%s %s %s %s")
        fname, action = v[0], v[1]
        if "ACTIVE" in action and "PASSIVE" not in action:
            raise NotImplementedError("active molecules are outside this engine's scope")
        mol = read_molecule(os.path.join(pyrh_path, "rh", "Molecules", fname), elements)
        if n == 0 and "H2" not in mol.ID:
            raise RHInputError("First molecule should be H2")
        mols.append(mol)
    return mols


def molecular_line_range(pyrh_path, mol: Molecule):
    """(lambda_min, lambda_max, qwing) over a passive molecule's line lists,
    or None.  Used only to detect wavelengths where background molecular
    lines would contribute (rh/opacity.c:771-790)."""
    lo, hi, qw = math.inf, -math.inf, 0.0
    for f in mol.line_files:
        p = os.path.join(pyrh_path, "rh", "Molecules", f)
        if not os.path.exists(p):
            raise RHInputError(f"Unable to open inputfile {p}")
        L = list(data_lines(p))
        hdr = L[1].split()
        if len(hdr) >= 2:
            qw = max(qw, float(hdr[1]))
        for s in L[2:]:
            n, v = sscanf(s, "%lf")
            if n == 1:
                lo, hi = min(lo, v[0]), max(hi, v[0])
    if lo > hi:
        return None
    return lo, hi, qw


# ---------------------------------------------------------------------------
# Barklem (ABO) broadening tables (rh/barklem.c:58-130)
# ---------------------------------------------------------------------------

BARKLEM_DIMS = {"SP": (21, 18, 1.0, 1.3), "PD": (18, 18, 1.3, 2.3),
                "DF": (18, 18, 2.3, 3.3)}
BARKLEM_DELTA_NEFF = 0.1


@dataclass
class BarklemTable:
    neff1: np.ndarray
    neff2: np.ndarray
    cross: np.ndarray   # [N1, N2]
    alpha: np.ndarray


def read_barklem(path, kind) -> BarklemTable:
    n1, n2, e1, e2 = BARKLEM_DIMS[kind]
    if not os.path.exists(path):
        raise RHInputError(f"Unable to open input file {path}")
    with open(path) as f:
        lines = f.readlines()
    # 3 header lines, N1*N2 cross values, then 2 lines (the rest of the
    # current line and one header), then N1*N2 alpha values.
    body = "".join(lines[3:])
    toks = body.split()
    vals = []
    idx = 0
    while len(vals) < n1 * n2:
        vals.append(float(toks[idx]))
        idx += 1
    cross = np.array(vals).reshape(n1, n2)
    # find the position of the second block: fgets twice after the last
    # number read; emulate by locating the remaining text line-wise
    consumed = 0
    text = body
    pos = 0
    for _ in range(n1 * n2):
        while text[pos].isspace():
            pos += 1
        while pos < len(text) and not text[pos].isspace():
            pos += 1
    rest = text[pos:]
    # two fgets: remainder of current line, then one full line
    nl1 = rest.find("\n")
    rest = rest[nl1 + 1:]
    nl2 = rest.find("\n")
    rest = rest[nl2 + 1:]
    toks2 = rest.split()
    alpha = np.array([float(t) for t in toks2[:n1 * n2]]).reshape(n1, n2)
    neff1 = np.array([e1 + n * BARKLEM_DELTA_NEFF for n in range(n1)])
    neff2 = np.array([e2 + n * BARKLEM_DELTA_NEFF for n in range(n2)])
    return BarklemTable(neff1, neff2, cross, alpha)
