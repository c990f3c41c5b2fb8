"""Host model of the NLTE problem: one active atom solved by accelerated
lambda iteration under pyrh's rhf1d() (rh/rhf1d/pyrh_compute1dray.c:280-352).

build_nlte(setup, wave) restates the host-side set-up of the reference:
- the active atom's line and continuum wavelength grids
  (rh/getlambda.c:42-180) merged with the reference and output wavelengths
  (rh/sortlambda.c:42-230);
- each transition's span of the merged grid and its quadrature weights
  (rh/getlambda.c:185-220);
- explicit cross-sections spline-interpolated onto that span
  (rh/sortlambda.c:334-362);
- the active set of transitions at every wavelength, continua before lines
  (rh/sortlambda.c:290-420);
- the broadening constants of rh/broad.c:34-316;
- the collisional-rate tables of rh/collision.c:450-760;
- the Gauss-Legendre rays of rh/gaussleg.c:30-64;
- the iteration keywords.

The background (continuum) model is pyrh_amd.model.build_model on the same
merged grid, without the active atom's continua (rh/metal.c:90).  Anything
the device code does not implement raises NotImplementedError.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import constants as K
from .model import (RunSetup, _abo_scale, build_model, cubeconvol, gaunt_bf, linear, locate,
                    spline_coef, spline_eval)
from .rhdata import BARKLEM_DELTA_NEFF, read_barklem

TR_LINE, TR_CONT = 0, 1
VDW_NONE, VDW_UNSOLD, VDW_BARKLEM = 0, 1, 2

ORBITALS = "SPDFGHIJKLMNOQRTUV"


def _determinate(label, g):
    """determinate (rh/zeeman.c:37-84): (S, L, J) of a level from its label,
    or None when the label shows no parity or the level is composite."""
    p = len(label) - 1
    while p > 0 and label[p] not in "EO":
        p -= 1
    if p <= 0:
        return None
    words = label[:p + 1].split()
    last = words[-1][-3:]
    if not last[0].isdigit():
        return None
    S = (int(last[0]) - 1) / 2.0
    L = ORBITALS.index(last[1]) if last[1] in ORBITALS else -1
    J = (g - 1.0) / 2.0
    if J > L + S:
        return None
    return S, L, J


def _barklem_active(setup, atom, ln):
    """getBarklemactivecross (rh/barklem.c:202-313): the ABO cross-section
    and velocity exponent of an active line from the Barklem tables, or None
    (ionised lower level, undetermined quantum numbers, no table for the
    orbital pair, or effective quantum numbers off the table): the
    reference then warns and uses UNSOLD (rh/readatom.c:311-318)."""
    i, j = ln.i, ln.j
    if atom.stage[i] > 0:
        return None
    if atom.abo_level[i] != -1 and atom.abo_level[j] != -1:
        Ll, Lu = atom.abo_level[i], atom.abo_level[j]
    else:
        di, dj = _determinate(atom.label[i], atom.g[i]), _determinate(atom.label[j], atom.g[j])
        if di is None or dj is None:
            return None
        Ll, Lu = di[1], dj[1]
    pair = {(0, 1): "SP", (1, 0): "SP", (1, 2): "PD", (2, 1): "PD", (2, 3): "DF", (3, 2): "DF"}
    kind = pair.get((Ll, Lu))
    if kind is None:
        return None
    bs = read_barklem(setup.barklem_paths[kind], kind)
    Z = atom.stage[j] + 1
    ic = j + 1
    while atom.stage[ic] < atom.stage[j] + 1:
        ic += 1
    E_Ryd = K.E_RYDBERG / (1.0 + K.M_ELECTRON / (atom.weight * K.AMU))
    neff1 = Z * math.sqrt(E_Ryd / (atom.E[ic] - atom.E[i]))
    neff2 = Z * math.sqrt(E_Ryd / (atom.E[ic] - atom.E[j]))
    if Ll > Lu:
        neff1, neff2 = neff2, neff1
    if neff1 < bs.neff1[0] or neff1 > bs.neff1[-1]:
        return None
    f1 = locate(bs.neff1, neff1)
    f1 = f1 + (neff1 - bs.neff1[f1]) / BARKLEM_DELTA_NEFF
    if neff2 < bs.neff2[0] or neff2 > bs.neff2[-1]:
        return None
    f2 = locate(bs.neff2, neff2)
    f2 = f2 + (neff2 - bs.neff2[f2]) / BARKLEM_DELTA_NEFF
    N1, N2 = bs.cross.shape
    cross = cubeconvol(N2, N1, bs.cross, f2, f1)
    alpha = cubeconvol(N2, N1, bs.alpha, f2, f1)
    return _abo_scale(cross, alpha, atom.weight, setup.atoms[0].weight), alpha
STARK_NONE, STARK_ABS, STARK_QUAD = 0, 1, 2
COLL_TYPES = {"OMEGA": 0, "CE": 1, "CI": 2, "CP": 3, "CH": 4, "CH0": 5, "CH+": 6}
# the active set at one wavelength holds at most this many transitions
# (rh/sortlambda.c:259, N_MAX_OVERLAP)
MAX_OVERLAP = 100


def gauss_legendre(n):
    """GaussLeg(0, 1, n) of rh/gaussleg.c:30-64."""
    x = np.zeros(n)
    w = np.zeros(n)
    m = (n + 1) // 2
    xm, xl = 0.5 * (1.0 + 0.0), 0.5 * (1.0 - 0.0)
    for i in range(m):
        zz = math.cos(K.PI * (i + 0.75) / (n + 0.5))
        while True:
            p1, p2 = 1.0, 0.0
            for j in range(1, n + 1):
                p3 = p2
                p2 = p1
                p1 = (2.0 * (j - 0.5) * zz * p2 - (j - 1.0) * p3) / j
            pp = n * (zz * p1 - p2) / (zz * zz - 1.0)
            z1 = zz
            zz = z1 - p1 / pp
            if abs(p1 / pp) <= 3.0E-14:
                break
        x[i] = xm - xl * zz
        w[i] = 2.0 * xl / ((1.0 - zz * zz) * pp * pp)
        x[n - 1 - i] = xm + xl * zz
        w[n - 1 - i] = w[i]
    return x, w


def line_grid(ln, symmetric, vmicro_char):
    """getLambda (rh/getlambda.c:42-160) for a line without Zeeman
    extension (B_STRENGTH_CHAR = 0), one component: wavelengths [nm]."""
    if ln.qcore <= 0.0 or ln.qwing <= 0.0:
        raise ValueError("qcore or qwing not positive")
    if symmetric:
        Nl = ln.Nlambda if ln.Nlambda % 2 else ln.Nlambda + 1
    else:
        Nl = (ln.Nlambda // 2) if ln.Nlambda % 2 else (ln.Nlambda + 1) // 2
    beta = 1.0 if ln.qwing <= 2.0 * ln.qcore else ln.qwing / (2.0 * ln.qcore)
    y = beta + math.sqrt(beta * beta + (beta - 1.0) * Nl + 2.0 - 3.0 * beta)
    b = 2.0 * math.log(y) / (Nl - 1)
    a = ln.qwing / (Nl - 2.0 + y * y)
    q = [a * (la + (math.exp(b * la) - 1.0)) for la in range(Nl)]
    q_to_lambda = ln.lambda0 * (vmicro_char / K.CLIGHT)
    if symmetric:
        return np.array([ln.lambda0 + q_to_lambda * q[la] for la in range(Nl)])
    out = np.zeros(2 * Nl - 1)
    mid = Nl - 1
    out[mid] = ln.lambda0
    for la in range(1, Nl):
        d = q_to_lambda * q[la]
        out[mid - la] = ln.lambda0 - d
        out[mid + la] = ln.lambda0 + d
    return out


def _hunt(grid, v):
    """Index of an exact grid value (Hunt on the merged grid)."""
    i = int(np.searchsorted(grid, v))
    if i >= grid.size or grid[i] != v:
        raise RuntimeError("wavelength missing from the merged grid")
    return i


def rh_hunt(arr, value, ilow):
    """Hunt (rh/hunt.c:17-80) on an ascending array from the guess ilow:
    the index j with arr[j] <= value < arr[j + 1] as that routine finds it.
    Where the search hunts down onto a value equal to a grid point, the
    point becomes the upper bracket and the result is one index short; the
    reference's transition spans (rh/sortlambda.c:306-311, 380-385) carry
    that quirk, so the host model keeps it."""
    n = len(arr)
    if ilow <= 0 or ilow > n - 1:
        ilow, ihigh = 0, n
    else:
        inc = 1
        if value >= arr[ilow]:
            if ilow == n - 1:
                return ilow
            ihigh = ilow + inc
            while value >= arr[ihigh]:
                ilow = ihigh
                inc += inc
                ihigh = ilow + inc
                if ihigh >= n:
                    ihigh = n
                    break
        else:
            ihigh = ilow
            if ilow == 0:
                return ilow
            while value <= arr[ilow]:
                ihigh = ilow
                inc += inc
                ilow = ihigh - inc
                if ilow <= 0:
                    ilow = 0
                    break
    while ihigh - ilow > 1:
        mid = (ihigh + ilow) >> 1
        if value >= arr[mid]:
            ilow = mid
        else:
            ihigh = mid
    return ilow


@dataclass
class NLTEModel:
    lte: object                 # background model on the merged grid
    lam: np.ndarray             # merged grid [nm]
    ref_index: int
    atom_index: int             # index of the active atom in setup.atoms
    ID: str
    nlevel: int
    E: np.ndarray
    g: np.ndarray
    stage: np.ndarray
    abundance: float
    weight: float
    # transitions, continua first then lines (their active-set order)
    tr_type: np.ndarray
    tr_i: np.ndarray
    tr_j: np.ndarray
    tr_nblue: np.ndarray
    tr_nlam: np.ndarray
    tr_ptr: np.ndarray          # into tr_wlam / tr_alpha
    tr_wlam: np.ndarray
    tr_alpha: np.ndarray
    tr_line: np.ndarray         # line index of a line transition, else -1
    ln_lambda0: np.ndarray
    ln_Aji: np.ndarray
    ln_Bji: np.ndarray
    ln_Bij: np.ndarray
    ln_Grad: np.ndarray
    ln_vdw: np.ndarray
    ln_vdw_cross: np.ndarray
    ln_vdw_alpha: np.ndarray
    ln_vdw_cross_he: np.ndarray
    ln_stark: np.ndarray
    ln_stark_c: np.ndarray      # |cStark| (ABS) or cStark23 (QUAD)
    ln_stark_C: np.ndarray
    ln_stark_Cm: np.ndarray
    as_ptr: np.ndarray          # active sets per wavelength (CSR over lam)
    as_tr: np.ndarray
    as_bezier: np.ndarray       # 1: angle-dependent (scalar Bezier), 0: Feautrier
    coll_type: np.ndarray
    coll_i: np.ndarray
    coll_j: np.ndarray
    coll_ptr: np.ndarray        # into coll_T / coll_c / coll_M
    coll_T: np.ndarray
    coll_c: np.ndarray
    coll_M: np.ndarray          # spline second derivatives (0: linear table)
    muz: np.ndarray
    wmu: np.ndarray
    ray_H: np.ndarray           # Rayleigh cross-sections per wavelength [m^2]:
    ray_He: np.ndarray          # H ground level, He ground level, H2
    ray_H2: np.ndarray
    sigma_thomson: float
    he_atom: int                # index of the He model atom (-1: none)
    nmax_iter: int
    nmax_scatter: int
    iter_limit: float
    ng_order: int
    ng_delay: int
    ng_period: int
    isum: int
    out_index: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))


def _rayleigh_atom(atom, lam, vchar):
    """sigma_Rayleigh(lambda) of rh/rayleigh.c:35-100 (times the ground-level
    population gives the scattering opacity); zero where it is off."""
    out = np.zeros(lam.size)
    ground = [ln for ln in atom.lines if ln.i == 0]
    if atom.stage[0] != 0 or not ground:
        return out
    limit = min(ln.lambda0 * (1.0 + ln.qwing * vchar / K.CLIGHT) for ln in ground)
    C = 2 * K.PI * (K.Q_ELECTRON / K.EPSILON_0) * (K.Q_ELECTRON / K.M_ELECTRON) / K.CLIGHT
    sigma_e = 8.0 * K.PI / 3.0 * math.pow(
        K.Q_ELECTRON / (math.sqrt(4.0 * K.PI * K.EPSILON_0) * (math.sqrt(K.M_ELECTRON) * K.CLIGHT)), 4)
    for n, x in enumerate(lam):
        if x <= limit:
            continue
        fomega = 0.0
        for ln in atom.lines:
            red = ln.lambda0 * (1.0 + ln.qwing * vchar / K.CLIGHT)
            if ln.i == 0 and x > red:
                r = x / ln.lambda0
                lambda2 = 1.0 / (r * r - 1.0)
                l0 = ln.lambda0 * K.NM_TO_M
                f = ln.Aji * (atom.g[ln.j] / atom.g[0]) * (l0 * l0) / C
                fomega += f * (lambda2 * lambda2)
        out[n] = sigma_e * fomega
    return out


_RH2_LAM = [121.57, 130.00, 140.00, 150.00, 160.00, 170.00, 185.46, 186.27, 193.58, 199.05,
            230.29, 237.91, 253.56, 275.36, 296.81, 334.24, 404.77, 407.90, 435.96, 546.23,
            632.80]
_RH2_SIG = [2.35E-06, 1.22E-06, 6.80E-07, 4.24E-07, 2.84E-07, 2.00E-07, 1.25E-07, 1.22E-07,
            1.00E-07, 8.70E-08, 4.29E-08, 3.68E-08, 2.75E-08, 1.89E-08, 1.36E-08, 8.11E-09,
            3.60E-09, 3.48E-09, 2.64E-09, 1.04E-09, 5.69E-10]


def _rayleigh_h2(lam):
    """H2 Rayleigh cross-section of rh/hydrogen.c:938-995 [m^2] (Victor &
    Dalgarno 1969 table, Tarafdar & Vardya 1973 fit redward of it)."""
    out = np.zeros(lam.size)
    a = (8.779E+01, 1.323E+06, 2.245E+10)
    for n, x in enumerate(lam):
        if x < 121.57:
            continue
        if x <= _RH2_LAM[-1]:
            s = linear(_RH2_LAM, _RH2_SIG, float(x))
        else:
            l2 = 1.0 / (x * x)
            s = (a[0] + (a[1] + a[2] * l2) * l2) * (l2 * l2)
        out[n] = s * K.MEGABARN_TO_M2
    return out


def _collisions(atom):
    """Parse the collisional data after the continua (rh/collision.c:
    484-640) into tables; interpolation in T happens per column."""
    types, I, J, ptr, Ts, cs, Ms = [], [], [], [0], [], [], []
    T = None
    for raw in atom.raw_lines:
        tok = raw.split()
        if not tok:
            continue
        key = tok[0]
        if key == "TEMP":
            n = int(tok[1])
            T = [float(x) for x in tok[2:2 + n]]
            if len(T) != n:
                raise ValueError("collision TEMP record is short")
            continue
        if key.startswith("END"):
            break
        if key not in COLL_TYPES:
            raise NotImplementedError(f"collision keyword {key} of {atom.ID}")
        if T is None:
            raise ValueError("collision record before TEMP")
        i1, i2 = int(tok[1]), int(tok[2])
        c = [float(x) for x in tok[3:3 + len(T)]]
        if len(c) != len(T):
            raise ValueError(f"collision record {key} {i1} {i2} is short")
        types.append(COLL_TYPES[key])
        I.append(min(i1, i2))
        J.append(max(i1, i2))
        Ts.extend(T)
        cs.extend(c)
        Ms.extend(spline_coef(T, c) if len(T) > 2 else [0.0] * len(T))
        ptr.append(len(Ts))
    return (np.array(types, np.int32), np.array(I, np.int32), np.array(J, np.int32),
            np.array(ptr, np.int32), np.array(Ts), np.array(cs), np.array(Ms))


def _broadening(setup, atom, ln):
    """Constants of rh/broad.c:34-316 for one active line."""
    j, i = ln.j, ln.i
    vdw, cross, alpha, cross_he = VDW_NONE, 0.0, 0.0, 0.0
    cvdw = list(ln.cvdWaals)
    barklem = None
    if "BARKLEM" in ln.vdw:
        barklem = _barklem_active(setup, atom, ln)
        if barklem is not None:
            # rh/barklem.c:304-311: He by UNSOLD with cvdWaals[2] = 1
            cvdw = [barklem[0], barklem[1], 1.0, 0.0]
    elif "UNSOLD" not in ln.vdw:
        raise NotImplementedError(f"van der Waals type {ln.vdw} for an active line")
    if cvdw[0] > 0.0 or cvdw[2] > 0.0:
        He_w, He_ab = float(setup.elements.weight[1]), float(setup.elements.abund[1])
        H_w = setup.atoms[0].weight
        fourPIeps0 = 4.0 * K.PI * K.EPSILON_0
        vrel35_He = math.pow(8.0 * K.KBOLTZMANN / (K.PI * K.AMU * atom.weight)
                             * (1.0 + atom.weight / He_w), 0.3)
        Z = int(atom.stage[j]) + 1
        ic = j + 1
        while atom.stage[ic] < atom.stage[j] + 1:
            ic += 1
        deltaR = ((K.E_RYDBERG / (atom.E[ic] - atom.E[j])) ** 2
                  - (K.E_RYDBERG / (atom.E[ic] - atom.E[i])) ** 2)
        C625 = math.pow(2.5 * (K.Q_ELECTRON * K.Q_ELECTRON / fourPIeps0) * (K.ABARH / fourPIeps0)
                        * 2 * K.PI * ((Z * K.RBOHR) * (Z * K.RBOHR)) / K.HPLANCK * deltaR, 0.4)
        vrel35_H = math.pow(8.0 * K.KBOLTZMANN / (K.PI * K.AMU * atom.weight)
                            * (1.0 + atom.weight / H_w), 0.3)
        if barklem is None:
            # UNSOLD (also the fallback of BARKLEM) sets cvdWaals[1] =
            # cvdWaals[3] = 0 (rh/readatom.c:311-320); rh/broad.c:87-100
            cross = 8.08 * (cvdw[0] * vrel35_H + cvdw[2] * He_ab * vrel35_He) * C625
            vdw = VDW_UNSOLD
        else:
            # rh/broad.c:125-136: cross_H T^((1 - alpha)/2) + cross_He T^0.3
            cross, alpha = cvdw[0], cvdw[1]
            cross_he = 8.08 * cvdw[2] * He_ab * vrel35_He * C625
            vdw = VDW_BARKLEM
    stark, c0, C, Cm = STARK_NONE, 0.0, 0.0, 0.0
    if ln.cStark != 0.0:
        if ln.cStark < 0.0:
            stark, c0 = STARK_ABS, abs(ln.cStark)
        else:
            stark = STARK_QUAD
            m_el = K.M_ELECTRON / K.AMU
            C = 8.0 * K.KBOLTZMANN / (K.PI * K.AMU * atom.weight)
            Cm = (math.pow(1.0 + atom.weight / m_el, 0.16666667)
                  + math.pow(1.0 + atom.weight / 28.0, 0.16666667))
            Z = int(atom.stage[i]) + 1
            ic = i + 1
            while ic < atom.E.size and atom.stage[ic] < atom.stage[i] + 1:
                ic += 1
            if ic >= atom.E.size or atom.stage[ic] == atom.stage[i]:
                raise ValueError(f"no overlying continuum for level {i}")
            E_Ryd = K.E_RYDBERG / (1.0 + K.M_ELECTRON / (atom.weight * K.AMU))
            neff_l = Z * math.sqrt(E_Ryd / (atom.E[ic] - atom.E[i]))
            neff_u = Z * math.sqrt(E_Ryd / (atom.E[ic] - atom.E[j]))
            C4 = ((K.Q_ELECTRON * K.Q_ELECTRON) / (4.0 * K.PI * K.EPSILON_0) * K.RBOHR
                  * (2.0 * K.PI * (K.RBOHR * K.RBOHR) / K.HPLANCK) / (18.0 * (Z * Z) * (Z * Z))
                  * ((neff_u * (5.0 * (neff_u * neff_u) + 1.0)) ** 2
                     - (neff_l * (5.0 * (neff_l * neff_l) + 1.0)) ** 2))
            c0 = 11.37 * math.pow(ln.cStark * C4, 0.66666667)
    return vdw, cross, stark, c0, C, Cm, alpha, cross_he


def build_nlte(setup: RunSetup, wave, fudge=None) -> NLTEModel:
    """The NLTE problem of run directory `setup` with output wavelengths
    `wave` [nm, vacuum]."""
    kw = setup.kw
    act = [ia for ia, a in enumerate(setup.atoms) if a.active]
    if len(act) != 1:
        raise NotImplementedError("exactly one active atom is supported")
    ia = act[0]
    atom = setup.atoms[ia]
    if any(m.active for m in setup.molecules if hasattr(m, "active")):
        raise NotImplementedError("active molecules")
    if atom.ID == "H ":
        raise NotImplementedError("active hydrogen (linear Stark, HYDROGEN_LTE)")
    if atom.Nfixed:
        raise NotImplementedError("fixed transitions")
    if kw.float("VMACRO_TRESH") != 0.0:
        raise NotImplementedError("VMACRO_TRESH > 0 (static columns in NLTE)")
    if kw.bool("ALLOW_PASSIVE_BB"):
        raise NotImplementedError("passive bound-bound lines with an active atom "
                                  "(set ALLOW_PASSIVE_BB = FALSE)")
    if kw.str("STOKES_MODE") != "NO_STOKES":
        raise NotImplementedError("polarized NLTE (STOKES_MODE must be NO_STOKES)")
    if kw.float("B_STRENGTH_CHAR") != 0.0:
        raise NotImplementedError("B_STRENGTH_CHAR (Zeeman-extended line grids)")
    if kw.bool("LIMIT_MEMORY") or kw.str("STARTING_J") != "NEW_J":
        raise NotImplementedError("LIMIT_MEMORY or STARTING_J other than NEW_J")
    if kw.str("SOLVE_NE") not in ("NONE", "FALSE"):
        raise NotImplementedError("SOLVE_NE")
    if atom.initial_solution != "LTE_POPULATIONS":
        raise NotImplementedError(f"initial solution {atom.initial_solution}")
    for ln in atom.lines:
        if "VOIGT" not in ln.shape or "PRD" in ln.shape or ln.Ncomponent != 1:
            raise NotImplementedError(f"line shape {ln.shape} (only CRD VOIGT)")
    vchar = kw.float("VMICRO_CHAR") * K.KM_TO_M
    lam_ref = kw.float("LAMBDA_REF")

    # --- grids: every column is moving (VMACRO_TRESH = 0), so lines take
    #     full (asymmetric) profiles (rh/getlambda.c:66-71)
    lgrids = [line_grid(ln, False, vchar) for ln in atom.lines]
    cgrids = [np.asarray(c.lam, np.float64) for c in atom.conts]
    allw = ([lam_ref] if lam_ref > 0.0 else []) + [float(x) for x in wave]
    for gcur in cgrids:
        allw += gcur.tolist()
    for gcur in lgrids:
        allw += gcur.tolist()
    allw.sort()
    lam = [allw[0]]
    for x in allw[1:]:
        if x > lam[-1]:
            lam.append(x)
    lam = np.array(lam)
    nspect = lam.size
    ref_index = _hunt(lam, lam_ref) if lam_ref > 0.0 else -1

    # --- transitions (continua first, in file order, then lines)
    tr_type, tr_i, tr_j, tr_nb, tr_nl, tr_line = [], [], [], [], [], []
    wl_all, al_all, ptr = [], [], [0]
    # spans as SortLambda finds them: Nblue from a zero guess, Nred hunted
    # from the previous transition's Nred (rh/sortlambda.c:296-311, 376-385)
    nr = 0
    for c, gcur in zip(atom.conts, cgrids):
        nb, nr = rh_hunt(lam, gcur[0], 0), rh_hunt(lam, gcur[-1], nr)
        sub = lam[nb:nr + 1]
        if c.hydrogenic:
            Z = int(atom.stage[c.j])
            n_eff = Z * math.sqrt(K.E_RYDBERG / (atom.E[c.j] - atom.E[c.i]))
            gbf0 = gaunt_bf(c.lambda0, n_eff, Z)
            alpha = [c.alpha0 * gaunt_bf(x, n_eff, Z) / gbf0 * ((x / c.lambda0) ** 3)
                     for x in sub]
        else:
            M = spline_coef(list(c.lam), list(c.alpha))
            alpha = [spline_eval(list(c.lam), list(c.alpha), M, float(x)) for x in sub]
        n = sub.size
        for la in range(n):
            if la == 0:
                w = 0.5 * (sub[1] - sub[0])
            elif la == n - 1:
                w = 0.5 * (sub[la] - sub[la - 1])
            else:
                w = 0.5 * (sub[la + 1] - sub[la - 1])
            wl_all.append(w)
        al_all += alpha
        ptr.append(len(wl_all))
        tr_type.append(TR_CONT); tr_i.append(c.i); tr_j.append(c.j)
        tr_nb.append(nb); tr_nl.append(n); tr_line.append(-1)
    for kr, (ln, gcur) in enumerate(zip(atom.lines, lgrids)):
        nb, nr = rh_hunt(lam, gcur[0], 0), rh_hunt(lam, gcur[-1], nr)
        sub = lam[nb:nr + 1]
        n = sub.size
        dop = K.CLIGHT / ln.lambda0          # asymmetric profile: no factor 2
        for la in range(n):
            if la == 0:
                w = 0.5 * (sub[1] - sub[0])
            elif la == n - 1:
                w = 0.5 * (sub[la] - sub[la - 1])
            else:
                w = 0.5 * (sub[la + 1] - sub[la - 1])
            wl_all.append(w * dop)
        al_all += [0.0] * n
        ptr.append(len(wl_all))
        tr_type.append(TR_LINE); tr_i.append(ln.i); tr_j.append(ln.j)
        tr_nb.append(nb); tr_nl.append(n); tr_line.append(kr)
    ntr = len(tr_type)

    # --- active sets
    as_lists = [[] for _ in range(nspect)]
    for t in range(ntr):
        for ns in range(tr_nb[t], tr_nb[t] + tr_nl[t]):
            as_lists[ns].append(t)
    if max(len(x) for x in as_lists) >= MAX_OVERLAP:
        raise ValueError("too many overlapping transitions")
    as_ptr = np.zeros(nspect + 1, np.int32)
    for ns in range(nspect):
        as_ptr[ns + 1] = as_ptr[ns] + len(as_lists[ns])
    as_tr = np.array([t for x in as_lists for t in x], np.int32)
    # scalar Bezier where a (polarizable, hence angle-dependent) line is
    # active or the column moves with a line present; Feautrier elsewhere
    # (rh/rhf1d/formal.c:78-103 with no background lines)
    as_bez = np.array([1 if any(tr_type[t] == TR_LINE for t in x) else 0 for x in as_lists],
                      np.int32)

    # --- line constants and broadening
    nl = len(atom.lines)
    br = [_broadening(setup, atom, ln) for ln in atom.lines]
    ct = _collisions(atom)
    muz, wmu = gauss_legendre(kw.int("NRAYS"))
    he = [k for k, a in enumerate(setup.atoms) if a.ID == "HE"]
    he_atom = he[0] if he else -1
    sigma_T = 8.0 * K.PI / 3.0 * math.pow(
        K.Q_ELECTRON / (math.sqrt(4.0 * K.PI * K.EPSILON_0) * (math.sqrt(K.M_ELECTRON) * K.CLIGHT)), 4)
    wave_bg = [x for x in lam if x != lam_ref]
    lte = build_model(setup, np.array(wave_bg), nlte=True, fudge=fudge)
    if not np.array_equal(lte.lam, lam):
        raise RuntimeError("background grid differs from the merged grid")
    isum = kw.int("I_SUM")
    if isum < -1 or isum >= atom.E.size:
        raise ValueError(f"Invalid value for ISUM = {isum}")
    out_index = np.array([i for i in range(nspect) if lam[i] != lam_ref], np.int32)
    return NLTEModel(
        lte=lte, lam=lam, ref_index=ref_index, atom_index=ia, ID=atom.ID,
        nlevel=int(atom.E.size), E=np.asarray(atom.E), g=np.asarray(atom.g),
        stage=np.asarray(atom.stage, np.int32), abundance=float(atom.abundance),
        weight=float(atom.weight),
        tr_type=np.array(tr_type, np.int32), tr_i=np.array(tr_i, np.int32),
        tr_j=np.array(tr_j, np.int32), tr_nblue=np.array(tr_nb, np.int32),
        tr_nlam=np.array(tr_nl, np.int32), tr_ptr=np.array(ptr, np.int32),
        tr_wlam=np.array(wl_all), tr_alpha=np.array(al_all), tr_line=np.array(tr_line, np.int32),
        ln_lambda0=np.array([ln.lambda0 for ln in atom.lines]),
        ln_Aji=np.array([ln.Aji for ln in atom.lines]),
        ln_Bji=np.array([ln.Bji for ln in atom.lines]),
        ln_Bij=np.array([ln.Bij for ln in atom.lines]),
        ln_Grad=np.array([ln.Grad for ln in atom.lines]),
        ln_vdw=np.array([b[0] for b in br], np.int32),
        ln_vdw_cross=np.array([b[1] for b in br]),
        ln_vdw_alpha=np.array([b[6] for b in br]),
        ln_vdw_cross_he=np.array([b[7] for b in br]),
        ln_stark=np.array([b[2] for b in br], np.int32),
        ln_stark_c=np.array([b[3] for b in br]),
        ln_stark_C=np.array([b[4] for b in br]),
        ln_stark_Cm=np.array([b[5] for b in br]),
        as_ptr=as_ptr, as_tr=as_tr, as_bezier=as_bez,
        coll_type=ct[0], coll_i=ct[1], coll_j=ct[2], coll_ptr=ct[3], coll_T=ct[4],
        coll_c=ct[5], coll_M=ct[6], muz=muz, wmu=wmu,
        ray_H=_rayleigh_atom(setup.atoms[0], lam, vchar),
        ray_He=(_rayleigh_atom(setup.atoms[he_atom], lam, vchar) if he_atom >= 0
                else np.zeros(lam.size)),
        ray_H2=_rayleigh_h2(lam), sigma_thomson=sigma_T, he_atom=he_atom,
        nmax_iter=kw.int("N_MAX_ITER"), nmax_scatter=kw.int("N_MAX_SCATTER"),
        iter_limit=kw.float("ITER_LIMIT"), ng_order=kw.int("NG_ORDER"),
        ng_delay=kw.int("NG_DELAY"), ng_period=kw.int("NG_PERIOD"), isum=isum,
        out_index=out_index)
