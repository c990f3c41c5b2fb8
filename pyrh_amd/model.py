"""Host-side model: everything about one run directory and wavelength grid
that does not depend on the atmosphere column.

``RunSetup`` parses the inputs once (the reference re-parses them on every
``rhf1d`` call, rh/rhf1d/pyrh_compute1dray.c:130-301).  ``build_model``
turns a setup plus a wavelength list into flat arrays: per-wavelength
constants, the Kurucz lines that contribute at each wavelength, the
bound-free continua active at each wavelength with their cross-sections,
and the data the per-column equilibria need.  The arrays are handed to
the GPU library (and to the test oracle) through ``rhg_model_t``.

Per-wavelength quantities are evaluated here in the same order of
floating-point operations as the reference routines that compute them
per call, so they are bit-identical to the reference's values.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import constants as K
from . import opacity_tables as T
from .cscan import data_lines, sscanf
from .rhdata import (ELEMENTS, RHInputError, read_abundance, read_atoms_input,
                     read_barklem, read_keywords, read_molecules_input,
                     molecular_line_range, BARKLEM_DELTA_NEFF)

Q_WING = 20.0
RLK_LABEL_LENGTH = 10


# ---------------------------------------------------------------------------
# small numerical helpers restating reference utilities
# ---------------------------------------------------------------------------

def locate(arr, value):
    """Bisection of rh/hunt.c:92-117 (ascending or descending)."""
    n = len(arr)
    ilow, ihigh = 0, n
    if arr[n - 1] > arr[0]:
        while ihigh - ilow > 1:
            idx = (ihigh + ilow) >> 1
            if value >= arr[idx]:
                ilow = idx
            else:
                ihigh = idx
    else:
        while ihigh - ilow > 1:
            idx = (ihigh + ilow) >> 1
            if value <= arr[idx]:
                ilow = idx
            else:
                ihigh = idx
    return ilow


def spline_coef(x, y):
    """Second-derivative table of rh/spline.c:30-66 (natural spline)."""
    n = len(x)
    q = [0.0] * n
    u = [0.0] * n
    M = [0.0] * n
    hj = x[1] - x[0]
    D = (y[1] - y[0]) / hj
    for j in range(1, n - 1):
        hj1 = x[j + 1] - x[j]
        mu = hj / (hj + hj1)
        D1 = (y[j + 1] - y[j]) / hj1
        p = mu * q[j - 1] + 2
        q[j] = (mu - 1) / p
        u[j] = ((D1 - D) * 6 / (hj + hj1) - mu * u[j - 1]) / p
        hj = hj1
        D = D1
    M[n - 1] = 0.0
    for j in range(n - 2, -1, -1):
        M[j] = q[j] * M[j + 1] + u[j]
    return M


def spline_eval(x, y, M, xv):
    """rh/spline.c:70-97 with Locate (hunt=FALSE)."""
    asc = x[1] > x[0]
    xmin = x[0] if asc else x[-1]
    xmax = x[-1] if asc else x[0]
    if xv <= xmin:
        return y[0] if asc else y[-1]
    if xv >= xmax:
        return y[-1] if asc else y[0]
    j = locate(x, xv)
    hj = x[j + 1] - x[j]
    fx = (xv - x[j]) / hj
    fx1 = 1 - fx
    return (fx1 * y[j] + fx * y[j + 1]
            + (fx1 * (fx1 * fx1 - 1) * M[j] + fx * (fx * fx - 1) * M[j + 1]) * (hj * hj) / 6.0)


def linear(xt, yt, xv):
    """rh/linear.c:22-54 for one abscissa."""
    asc = xt[1] > xt[0]
    xmin = xt[0] if asc else xt[-1]
    xmax = xt[-1] if asc else xt[0]
    if xv <= xmin:
        return yt[0] if asc else yt[-1]
    if xv >= xmax:
        return yt[-1] if asc else yt[0]
    j = locate(xt, xv)
    fx = (xt[j + 1] - xv) / (xt[j + 1] - xt[j])
    return fx * yt[j] + (1 - fx) * yt[j + 1]


def gaunt_bf(lam, n_eff, charge):
    """M. J. Seaton (1960) bound-free Gaunt factor, rh/hydrogen.c:246-259."""
    x = ((K.HPLANCK * K.CLIGHT) / (lam * K.NM_TO_M)) / (K.E_RYDBERG * (charge * charge))
    x3 = math.pow(x, 0.33333333)
    nsqx = 1.0 / ((n_eff * n_eff) * x)
    return (1.0 + 0.1728 * x3 * (1.0 - 2.0 * nsqx)
            - 0.0496 * (x3 * x3) * (1.0 - (1.0 - nsqx) * 0.66666667 * nsqx))


def frac_index(table, value):
    """index + fraction as the ff routines form it after Hunt."""
    i = locate(table, value)
    if i == len(table) - 1:
        return float(i)   # value equals the last node: zero-weight step
    return i + (value - table[i]) / (table[i + 1] - table[i])


def gammln(xx):
    """ln Gamma by the Lanczos series of Numerical Recipes (Press et al.
    1986), as rh/gammafunc.c:23-40 uses for the ABO cross-sections."""
    cof = (76.18009172947146, -86.50532032941677, 24.01409824083091,
           -1.231739572450155, 0.1208650973866179e-2, -0.5395239384953e-5)
    x = y = xx
    tmp = x + 5.5
    tmp -= (x + 0.5) * math.log(tmp)
    ser = 1.000000000190015
    for c in cof:
        y += 1
        ser += c / y
    return -tmp + math.log(2.5066282746310005 * ser / x)


def cc_kernel(s):
    """Keys (1981) cubic-convolution kernel (rh/cubeconvol.c:150-160)."""
    return [-s * (1.0 - s * (2.0 - s)) / 2.0,
            1.0 + s * s * (3.0 * s - 5.0) / 2.0,
            s * (1.0 + s * (4.0 - 3.0 * s)) / 2.0,
            s * s * (s - 1.0) / 2.0]


def cubeconvol(Nx, Ny, f, x, y):
    """2-D cubic convolution on the fractional indices (x, y) of the
    row-major table f[Ny][Nx], with the boundary extrapolation of
    rh/cubeconvol.c:30-140."""
    f = np.asarray(f, dtype=np.float64).ravel()
    if x <= 0.0:
        i, ux = 0, cc_kernel(0.0)
    elif x >= Nx - 1:
        i, ux = Nx - 2, cc_kernel(1.0)
    else:
        ip = math.floor(x)
        i, ux = int(ip), cc_kernel(x - ip)
    if y <= 0.0:
        j, uy = 0, cc_kernel(0.0)
    elif y >= Ny - 1:
        j, uy = Ny - 2, cc_kernel(1.0)
    else:
        jp = math.floor(y)
        j, uy = int(jp), cc_kernel(y - jp)

    def F(row, col):
        return f[Nx * row + col]

    base_r, base_c = j - 1, i - 1
    c = [[0.0] * 4 for _ in range(4)]
    interior_r = not (j == 0 or j == Ny - 2)
    interior_c = not (i == 0 or i == Nx - 2)
    if interior_r and interior_c:
        g = 0.0
        for m in range(4):
            for n in range(4):
                g += F(base_r + m, base_c + n) * ux[n] * uy[m]
        return g
    mrange = range(1, 4) if j == 0 else (range(0, 3) if j == Ny - 2 else range(4))
    for m in mrange:
        if i == 0:
            for n in range(1, 4):
                c[m][n] = F(base_r + m, base_c + n)
            c[m][0] = 3.0 * (c[m][1] - c[m][2]) + c[m][3]
        elif i == Nx - 2:
            for n in range(0, 3):
                c[m][n] = F(base_r + m, base_c + n)
            c[m][3] = 3.0 * (c[m][2] - c[m][1]) + c[m][0]
        else:
            for n in range(4):
                c[m][n] = F(base_r + m, base_c + n)
    if j == 0:
        for n in range(4):
            c[0][n] = 3.0 * (c[1][n] - c[2][n]) + c[3][n]
    elif j == Ny - 2:
        for n in range(4):
            c[3][n] = 3.0 * (c[2][n] - c[1][n]) + c[0][n]
    g = 0.0
    for m in range(4):
        for n in range(4):
            g += c[m][n] * ux[n] * uy[m]
    return g


# ---------------------------------------------------------------------------
# Kurucz line list (rh/kurucz.c:121-430)
# ---------------------------------------------------------------------------

ORBITALS = "SPDFGHIJKLMNOQRTUVWX"


def lande(S, L, J):
    """rh/zeeman.c:128-134."""
    if J == 0.0:
        return 0.0
    return 1.5 + (S * (S + 1.0) - L * (L + 1)) / (2.0 * J * (J + 1.0))


def zeeman_strength(Ju, Mu, Jl, Ml):
    """Relative strength of a Zeeman component, Stenflo (1994) p. 108;
    rh/zeeman.c:137-190."""
    q = int(Ml - Mu)
    dJ = int(Ju - Jl)
    if dJ == 0:
        return {0: 2.0 * Mu * Mu, -1: (Ju + Mu) * (Ju - Mu + 1.0),
                1: (Ju - Mu) * (Ju + Mu + 1.0)}[q]
    if dJ == 1:
        return {0: 2.0 * (Ju * Ju - Mu * Mu), -1: (Ju + Mu) * (Ju + Mu - 1.0),
                1: (Ju - Mu) * (Ju - Mu - 1.0)}[q]
    if dJ == -1:
        return {0: 2.0 * ((Ju + 1.0) * (Ju + 1.0) - Mu * Mu),
                -1: (Ju - Mu + 1.0) * (Ju - Mu + 2.0),
                1: (Ju + Mu + 1.0) * (Ju + Mu + 2.0)}[q]
    raise RHInputError(f"Invalid dJ: {dJ}")


def _label_SL(label):
    """RLKdeterminate (rh/kurucz.c:925-969): S and L from the last word of
    a 10-character Kurucz level label, or None."""
    words = label.split(" ")
    words = [w for w in words if w]
    if not words:
        return None
    w = words[-1]
    tail = w[-2:] if len(w) >= 2 else w
    n, v = sscanf(tail, "%d%1s")
    if n != 2 or not v[1].isupper():
        return None
    mult, orb = v
    if orb not in ORBITALS:
        raise RHInputError(f"Invalid orbital: {orb}")
    return (mult - 1) / 2.0, ORBITALS.index(orb)


@dataclass
class RLKLine:
    pt_index: int
    stage: int
    Ei: float
    Ej: float
    gi: float
    gj: float
    lambda0: float        # nm vacuum
    Aji: float
    Bji: float
    Bij: float
    Grad: float
    GStark: float
    GvdWaals: float
    vdwaals: int
    cross: float
    alpha: float
    isotope_frac: float
    hyperfine_frac: float
    polarizable: bool
    comps: list           # [(q, shift, strength)]
    line_index: int
    gL_i: float = 0.0
    gL_j: float = 0.0


def _rlk_records(path):
    """fgets(…, 172) chunks that do not start with '#' (rh/kurucz.c:168-176)."""
    out = []
    with open(path, "r", errors="replace") as f:
        for line in f:
            if len(line.rstrip("\n")) > 170:
                raise NotImplementedError(
                    f"Kurucz record longer than 170 characters in {path}")
            if line[:1] != "#":
                out.append(line)
    return out


def _barklem_cross(bs, rlk, element_ionpot, element_weight, H_weight):
    """getBarklemcross (rh/barklem.c:137-197) for a neutral line."""
    if rlk["stage"] > 0:
        return None
    dEi = element_ionpot - rlk["Ei"]
    if dEi <= 0.0:
        return None
    dEj = element_ionpot - rlk["Ej"]
    if dEj <= 0.0:
        return None
    Z = float(rlk["stage"] + 1)
    E_Ryd = K.E_RYDBERG / (1.0 + K.M_ELECTRON / (element_weight * K.AMU))
    neff1 = Z * math.sqrt(E_Ryd / dEi)
    neff2 = Z * math.sqrt(E_Ryd / dEj)
    if rlk["li"] > rlk["lj"]:
        neff1, neff2 = neff2, neff1
    if neff1 < bs.neff1[0] or neff1 > bs.neff1[-1]:
        return None
    idx = locate(bs.neff1, neff1)
    f1 = idx + (neff1 - bs.neff1[idx]) / BARKLEM_DELTA_NEFF
    if neff2 < bs.neff2[0] or neff2 > bs.neff2[-1]:
        return None
    idx = locate(bs.neff2, neff2)
    f2 = idx + (neff2 - bs.neff2[idx]) / BARKLEM_DELTA_NEFF
    N1, N2 = bs.cross.shape
    cross = cubeconvol(N2, N1, bs.cross, f2, f1)
    alpha = cubeconvol(N2, N1, bs.alpha, f2, f1)
    return cross, alpha


def _abo_scale(cross, alpha, element_weight, H_weight):
    reducedmass = K.AMU / (1.0 / H_weight + 1.0 / element_weight)
    meanvelocity = math.sqrt(8.0 * K.KBOLTZMANN / (K.PI * reducedmass))
    crossmean = (K.RBOHR * K.RBOHR) * math.pow(meanvelocity / 1.0E4, -alpha)
    return cross * (2.0 * math.pow(4.0 / K.PI, alpha / 2.0)
                    * math.exp(gammln((4.0 - alpha) / 2.0)) * meanvelocity * crossmean)


def read_kurucz_lines(list_file, cwd, setup) -> list:
    """readKuruczLines (rh/kurucz.c:121-430) for every file named in the
    KURUCZ_DATA list.  List entries are opened relative to the process
    working directory as the reference does (rh/kurucz.c:160-163); if not
    found there they are tried relative to the run directory."""
    if list_file in (None, "none") or not os.path.exists(list_file):
        if list_file not in (None, "none"):
            raise RHInputError(f"Unable to open input file {list_file}")
        return []
    el = setup.elements
    C = 2.0 * K.PI * (K.Q_ELECTRON / K.EPSILON_0) * (K.Q_ELECTRON / K.M_ELECTRON) / K.CLIGHT
    bs = {k: read_barklem(setup.barklem_paths[k], k) for k in ("SP", "PD", "DF")}
    H_weight = setup.atoms[0].weight
    He_abund = float(el.abund[1])
    lines = []
    for entry in data_lines(list_file):
        n, v = sscanf(entry, "%s")
        fname = v[0]
        path = fname if os.path.exists(fname) else os.path.join(cwd, fname)
        if not os.path.exists(path):
            raise RHInputError(f"Unable to open input file {fname}")
        line_index = 0
        for rec in _rlk_records(path):
            nr, v = sscanf(rec, "%lf %lf %s %lf")
            lambda_air, gf, elem_code, Ei = (v + [None] * 4)[:4]
            nread = max(nr, 0)
            _, pv = sscanf(elem_code, "%d.%d")
            pt_index, stage = pv[0], pv[1]
            n1, vv = sscanf(rec[53:], "%lf")
            Ej = vv[0] if n1 > 0 else 0.0
            nread += max(n1, 0)
            Ei = abs(Ei) * (K.HPLANCK * K.CLIGHT) / K.CM_TO_M
            Ej = abs(Ej) * (K.HPLANCK * K.CLIGHT) / K.CM_TO_M
            if Ej < Ei:
                swap = True
                rEi, rEj = Ej, Ei
                labeli, labelj = rec[69:69 + RLK_LABEL_LENGTH], rec[41:41 + RLK_LABEL_LENGTH]
            else:
                swap = False
                rEi, rEj = Ei, Ej
                labeli, labelj = rec[41:41 + RLK_LABEL_LENGTH], rec[69:69 + RLK_LABEL_LENGTH]
            n1, vv = sscanf(rec[35:], "%lf")
            Ji = vv[0]
            nread += max(n1, 0)
            n1, vv = sscanf(rec[63:], "%lf")
            Jj = vv[0]
            nread += max(n1, 0)
            if swap:
                Ji, Jj = Jj, Ji
            gi, gj = 2 * Ji + 1, 2 * Jj + 1
            lam_air = lambda_air
            if setup.lam_values is not None:
                for idl, lid in enumerate(setup.lam_ids):
                    if lid == line_index:
                        lam_air += setup.lam_values[idl]
            lambda0 = K.air_to_vacuum(lam_air) * K.NM_TO_M
            rEj = rEi + (K.HPLANCK * K.CLIGHT) / lambda0
            Aji = C / (lambda0 * lambda0) * K.pow10(gf) / gj
            if setup.loggf_values is not None:
                for idl, lid in enumerate(setup.loggf_ids):
                    if lid == line_index:
                        Aji = C / (lambda0 * lambda0) * K.pow10(setup.loggf_values[idl]) / gj
            Bji = (lambda0 * lambda0 * lambda0) / (2.0 * K.HPLANCK * K.CLIGHT) * Aji
            Bij = (gj / gi) * Bji
            rlk = dict(stage=stage, Ei=rEi, Ej=rEj, li=None, lj=None)
            got_abo = False
            cross = alpha = 0.0
            nl, lv = sscanf(rec[160:], "%lf %lf") if len(rec) > 160 else (-1, [])
            if nl != -1:
                a_, s_ = (lv + [None, None])[:2]
                if a_ in (0.0, 1.0, 2.0, 3.0) or s_ in (0.0, 1.0, 2.0, 3.0):
                    li, lj = int(a_), int(s_) if s_ is not None else None
                    if swap:
                        li, lj = lj, li
                    rlk["li"], rlk["lj"] = li, lj
                else:
                    got_abo = True
                    alpha, cross = a_, s_
            SLi = _label_SL(labeli)
            SLj = _label_SL(labelj) if SLi is not None else None
            determined = SLi is not None and SLj is not None
            polarizable = setup.stokes and determined
            G = rec[79:79 + 18]
            n3, gv = sscanf(G, "%lf %lf %lf")
            nread += max(n3, 0)
            if n3 < 3:
                raise RHInputError(f"Unable to read broadening of Kurucz line in {path}")
            Grad, GStark, GvdW = gv
            GStark = K.pow10(GStark) * K.CUBE_CM_TO_M if GStark != 0.0 else 0.0
            GvdWaals = K.pow10(GvdW) * K.CUBE_CM_TO_M if GvdW != 0.0 else 0.0
            elem = pt_index - 1
            ionpot = el.ionpot[elem][stage] if stage < len(el.ionpot[elem]) else None
            vdwaals = None
            if got_abo:
                cross = _abo_scale(cross, alpha, el.weight[elem], H_weight)
                vdwaals = 2
            else:
                # reference reads li/lj uninitialised when a record has no
                # l-numbers; such lines are treated as having none here
                pair = (rlk["li"], rlk["lj"])
                kind = ("SP" if pair in ((0, 1), (1, 0)) else
                        "PD" if pair in ((1, 2), (2, 1)) else
                        "DF" if pair in ((2, 3), (3, 2)) else None)
                if kind is not None and ionpot is not None:
                    r = _barklem_cross(bs[kind], rlk, ionpot, el.weight[elem], H_weight)
                    if r is not None:
                        cross = _abo_scale(r[0], r[1], el.weight[elem], H_weight)
                        alpha = r[1]
                        vdwaals = 2
            if vdwaals is None:
                # getUnsoldcross (rh/kurucz.c:985-1019)
                if stage > el.nstage[elem] - 1:
                    vdwaals = 0
                else:
                    Z = stage + 1
                    dR = ((K.E_RYDBERG / (ionpot - rEj)) ** 2
                          - (K.E_RYDBERG / (ionpot - rEi)) ** 2)
                    dR = ((K.E_RYDBERG / (ionpot - rEj)) * (K.E_RYDBERG / (ionpot - rEj))
                          - (K.E_RYDBERG / (ionpot - rEi)) * (K.E_RYDBERG / (ionpot - rEi)))
                    if dR <= 0.0:
                        vdwaals = 0
                    else:
                        w = el.weight[elem]
                        vrel35_H = math.pow(8.0 * K.KBOLTZMANN / (K.PI * K.AMU * w)
                                            * (1.0 + w / el.weight[0]), 0.3)
                        vrel35_He = math.pow(8.0 * K.KBOLTZMANN / (K.PI * K.AMU * w)
                                             * (1.0 + w / el.weight[1]), 0.3)
                        FOURPIEPS0 = 4.0 * K.PI * K.EPSILON_0
                        C625 = math.pow(2.5 * ((K.Q_ELECTRON * K.Q_ELECTRON) / FOURPIEPS0)
                                        * (K.ABARH / FOURPIEPS0)
                                        * 2 * K.PI * ((Z * K.RBOHR) * (Z * K.RBOHR)) / K.HPLANCK
                                        * dR, 0.4)
                        cross = 8.08 * (vrel35_H + He_abund * vrel35_He) * C625
                        vdwaals = 1
            Grad = K.pow10(Grad) if Grad != 0.0 else Aji
            n1, vv = sscanf(rec[106:], "%d")
            nread += max(n1, 0)
            n1, vv = sscanf(rec[108:], "%lf")
            iso = K.pow10(vv[0])
            nread += max(n1, 0)
            n1, vv = sscanf(rec[117:], "%lf")
            hfr = K.pow10(vv[0])
            nread += max(n1, 0)
            n1, vv = sscanf(rec[123:], "%5d%5d")
            nread += max(n1, 0)
            n1, vv = sscanf(rec[143:], "%5d%5d")
            nread += max(n1, 0)
            gL_i = vv[0] * K.MILLI
            gL_j = vv[1] * K.MILLI
            if swap:
                gL_i, gL_j = gL_j, gL_i
            if not setup.LS_lande and setup.stokes:
                if gL_i != -99 * K.MILLI and gL_j != -99 * K.MILLI:
                    polarizable = True
            if nread < 17:
                raise RHInputError(f"Unable to read Kurucz record in {path}")
            comps = []
            if polarizable:
                comps = _rlk_zeeman(gi, gj, SLi, SLj, gL_i, gL_j, setup.LS_lande)
            lines.append(RLKLine(pt_index, stage, rEi, rEj, gi, gj, lambda0 / K.NM_TO_M,
                                 Aji, Bji, Bij, Grad, GStark, GvdWaals, vdwaals, cross, alpha,
                                 iso, hfr, polarizable, comps, line_index, gL_i, gL_j))
            line_index += 1
    # qsort by lambda0 (rh/background.c:293); the sort is not stable, equal
    # wavelengths keep file order here
    lines.sort(key=lambda r: r.lambda0)
    return lines


def _rlk_zeeman(gi, gj, SLi, SLj, gL_i, gL_j, LS_lande):
    """RLKZeeman (rh/kurucz.c:832-920): components with strengths
    normalised per q, in the reference's (Ml, Mu) loop order."""
    Jl = (gi - 1.0) / 2.0
    Ju = (gj - 1.0) / 2.0
    if LS_lande:
        gLl = lande(SLi[0], SLi[1], Jl)
        gLu = lande(SLj[0], SLj[1], Ju)
    else:
        gLl = lande(SLi[0], SLi[1], Jl) if gL_i == -99 * K.MILLI else gL_i
        gLu = lande(SLj[0], SLj[1], Ju) if gL_j == -99 * K.MILLI else gL_j
    comps = []
    norm = [0.0, 0.0, 0.0]
    Ml = -Jl
    while Ml <= Jl:
        Mu = -Ju
        while Mu <= Ju:
            if abs(Mu - Ml) <= 1.0:
                q = int(Ml - Mu)
                s = zeeman_strength(Ju, Mu, Jl, Ml)
                comps.append([q, gLl * Ml - gLu * Mu, s])
                norm[q + 1] += s
            Mu += 1
        Ml += 1
    for c in comps:
        c[2] = c[2] / norm[c[0] + 1]
    return [tuple(c) for c in comps]


# ---------------------------------------------------------------------------
# Run setup (parsing once per run directory)
# ---------------------------------------------------------------------------

@dataclass
class RunSetup:
    cwd: str
    pyrh_path: str
    kw: object
    elements: object
    atoms: list
    molecules: list
    rlk: list
    barklem_paths: dict
    stokes: bool = True
    LS_lande: bool = True
    loggf_ids: object = None
    loggf_values: object = None
    lam_ids: object = None
    lam_values: object = None
    mol_ranges: list = field(default_factory=list)


def open_run(cwd, pyrh_path=None, atomic_number=None, atomic_abundance=None,
             loggf_ids=None, loggf_values=None, lam_ids=None, lam_values=None) -> RunSetup:
    """Parse keyword.input and everything it references, as the start of
    rhf1d() does (rh/rhf1d/pyrh_compute1dray.c:130-301)."""
    pyrh_path = pyrh_path or os.environ.get("PYRH_PATH")
    if not pyrh_path:
        raise RHInputError("PYRH_PATH is not set (needed for rh/Atoms and rh/Molecules)")
    kw = read_keywords(os.path.join(cwd, "keyword.input"))
    atoms_dir = os.path.join(pyrh_path, "rh", "Atoms")

    def dflt(key, name):
        v = kw.str(key)
        return v if v else os.path.join(atoms_dir, name)

    abund_path = dflt("ABUND_FILE", "abundance.input")
    pf_path = dflt("KURUCZ_PF_DATA", "pf_Kurucz.input")
    barklem = {"SP": dflt("BARKLEM_SP_DATA", "Barklem_spdata.dat"),
               "PD": dflt("BARKLEM_PD_DATA", "Barklem_pddata.dat"),
               "DF": dflt("BARKLEM_DF_DATA", "Barklem_dfdata.dat")}
    nab = 0 if atomic_number is None else len(atomic_number)
    elements = read_abundance(abund_path, pf_path, kw.float("METALLICITY"), nab,
                              atomic_number, atomic_abundance)
    atoms = read_atoms_input(os.path.join(cwd, kw.str("ATOMS_FILE")), pyrh_path, elements)
    mols = read_molecules_input(os.path.join(cwd, kw.str("MOLECULES_FILE")), pyrh_path,
                                elements)
    setup = RunSetup(cwd, pyrh_path, kw, elements, atoms, mols, [], barklem,
                     stokes=True, LS_lande=kw.bool("LS_LANDE"),
                     loggf_ids=loggf_ids, loggf_values=loggf_values,
                     lam_ids=lam_ids, lam_values=lam_values)
    kdata = kw.str("KURUCZ_DATA")
    if kdata != "none":
        setup.rlk = read_kurucz_lines(os.path.join(cwd, kdata), cwd, setup)
    setup.mol_ranges = [molecular_line_range(pyrh_path, m) if m.line_files else None
                        for m in mols]
    return setup


# ---------------------------------------------------------------------------
# Model (wavelength-grid dependent arrays)
# ---------------------------------------------------------------------------

STOKES_MODES = {"NO_STOKES": 0, "FIELD_FREE": 1, "POLARIZATION_FREE": 2, "FULL_STOKES": 3}


@dataclass
class Model:
    setup: RunSetup
    lam: np.ndarray          # [nspect] all wavelengths incl. reference
    ref_index: int
    out_index: np.ndarray    # indices of output wavelengths
    arrays: dict             # name -> numpy array (fields of rhg_model_t)
    scalars: dict
    cont_desc: list          # (atom index, continuum object) per continuum


def _check_supported(setup, nlte=False):
    kw = setup.kw
    if not kw.bool("HYDROGEN_LTE"):
        raise NotImplementedError("HYDROGEN_LTE = FALSE (NLTE hydrogen in the background)")
    if kw.str("S_INTERPOLATION") != "S_BEZIER3":
        raise NotImplementedError("only S_INTERPOLATION = S_BEZIER3 is implemented")
    if kw.str("S_INTERPOLATION_STOKES") != "DELO_BEZIER3":
        raise NotImplementedError("only S_INTERPOLATION_STOKES = DELO_BEZIER3 is implemented")
    if kw.bool("RLK_SCATTER"):
        raise NotImplementedError("RLK_SCATTER = TRUE")
    if kw.bool("BACKGROUND_POLARIZATION"):
        raise NotImplementedError("BACKGROUND_POLARIZATION = TRUE")
    if nlte and kw.str("ATMOS_ITOP") != "none":
        raise NotImplementedError("ATMOS_ITOP irradiation at the top boundary with an ACTIVE atom")
    if kw.str("STOKES_MODE") not in STOKES_MODES:
        raise RHInputError(f"Unknown STOKES_MODE = {kw.str('STOKES_MODE')}")
    if not nlte and any(a.active for a in setup.atoms):
        raise NotImplementedError("active atoms (NLTE) are not solved by the LTE path; "
                                  "use pyrh_amd.nlte")
    if kw.float("METALLICITY") != 0.0:
        pass  # handled in read_abundance
    if kw.bool("NON_ICE"):
        raise NotImplementedError("NON_ICE = TRUE")


def atomic_rf_lines(setup: RunSetup, loggf_ids=None, nlam=0) -> np.ndarray:
    """Model line (index into setup.rlk) of each atomic response-function
    parameter, in pyrh's order: the log gf ids, then the nlam wavelength-shift
    ids (pyrh.pyx:658-660).  -1 where the reference fills nothing: a line
    takes the last log gf id that names its file-local index
    (rh/kurucz.c:241-251), and the wavelength-shift entries are never filled
    (rh/kurucz.c:225-232)."""
    ids = [] if loggf_ids is None else [int(x) for x in loggf_ids]
    par = np.full(len(ids) + int(nlam), -1, np.int32)
    for l, r in enumerate(setup.rlk):
        last = None
        for idl, lid in enumerate(ids):
            if lid == r.line_index:
                last = idl
        if last is not None:
            if par[last] != -1:
                raise NotImplementedError(
                    f"log gf id {ids[last]} names a line in more than one Kurucz file; the "
                    "reference's response then depends on the order of its opacity loop")
            par[last] = l
    return par


def build_model(setup: RunSetup, wave, nlte=False, check=True, lam_ref=None,
                fudge=None) -> Model:
    """Wavelength grid and all column-independent tables for LTE synthesis.
    check=False skips the keyword checks of the synthesis (for the
    hydrostatic-equilibrium builder and the depth scales, which need only
    the background); lam_ref replaces LAMBDA_REF (get_scales).  fudge =
    (fudge_wave [n], fudge_value [3][n]) are pyrh's opacity fudge factors
    (rh/rhf1d/pyrh_compute1dray.c:183-195): row 0 scales the H- opacity
    (rh/background.c:362-369, the per-wavelength factor RHG_LC_HMIN_FUDGE),
    row 2 the metal bound-free opacity (background.c:438-451, folded into
    the metal continua's cross-sections), row 1 the scattering opacity
    (background.c:453-462, RHG_LC_SCATT_FUDGE; it enters only an NLTE
    solution), each by Linear in wavelength."""
    if check:
        _check_supported(setup, nlte)
    kw = setup.kw
    el = setup.elements
    atoms = setup.atoms
    lam_ref = kw.float("LAMBDA_REF") if lam_ref is None else float(lam_ref)
    vchar = kw.float("VMICRO_CHAR") * K.KM_TO_M
    vmacro = kw.float("VMACRO_TRESH") * K.KM_TO_M

    # --- SortLambda (rh/sortlambda.c:42-225): ref + table, sort, unique
    wl = list(map(float, wave))
    if lam_ref > 0.0:
        wl = [lam_ref] + wl
    wl.sort()
    lam = [wl[0]]
    for x in wl[1:]:
        if x > lam[-1]:
            lam.append(x)
    lam = np.array(lam)
    nspect = lam.size
    ref_index = locate(lam, lam_ref) if lam_ref > 0.0 else -1
    out_index = np.array([i for i in range(nspect) if lam[i] != lam_ref], np.int32)

    # --- per-wavelength constants
    hc_k = (K.HPLANCK * K.CLIGHT) / (K.KBOLTZMANN * K.NM_TO_M)
    twohc = (2.0 * K.HPLANCK * K.CLIGHT) / (K.NM_TO_M * K.NM_TO_M * K.NM_TO_M)
    hmbf_l = list(T.HMIN_BF_LAMBDA)
    hmbf_a = list(T.HMIN_BF_ALPHA)
    hmbf_M = spline_coef(hmbf_l, hmbf_a)
    NC = 23
    lc = np.zeros((nspect, NC))
    if fudge is not None:
        fw = np.ascontiguousarray(fudge[0], np.float64).ravel()
        fv = np.ascontiguousarray(fudge[1], np.float64)
        if fv.ndim != 2 or fv.shape[0] != 3 or fv.shape[1] != fw.size or fw.size < 2:
            raise ValueError("fudge_value must be [3][len(fudge_wave)] with at least two "
                             "wavelengths")
        hmin_f = np.array([linear(fw, fv[0], float(x)) for x in lam])
        scatt_f = np.array([linear(fw, fv[1], float(x)) for x in lam])
        metal_f = np.array([linear(fw, fv[2], float(x)) for x in lam])
    else:
        hmin_f = scatt_f = metal_f = np.ones(nspect)
    lc[:, 20] = hmin_f
    lc[:, 21] = scatt_f
    # irradiation at the top (ATMOS_ITOP, rh/rhf1d/multiatmos.c:185-222): an
    # XDR file of Nspect x Nrays doubles on the merged grid, read after
    # rhf1d() has set one ray (rh/rhf1d/pyrh_compute1dray.c:293, 308); the
    # Feautrier solver of the emergent ray uses it (feautrier.c:94-96), and
    # so does the inward scalar Bezier ray of the log gf response functions
    # (bezier_1D.c:370-380).  A relative path is taken from the run
    # directory; pyrh opens it from the process's working directory.
    itop = kw.str("ATMOS_ITOP")
    if itop != "none":
        path = itop if os.path.isabs(itop) else os.path.join(setup.cwd, itop)
        if not os.path.exists(path):
            raise RHInputError(f"Unable to open inputfile {itop}")
        v = np.fromfile(path, dtype=">f8")
        if v.size < nspect:
            raise RHInputError("Unable to read irradiation data at TOP of atmosphere")
        lc[:, 22] = v[:nspect]
    for n, la in enumerate(lam):
        lc[n, 0] = la
        lc[n, 1] = (K.HPLANCK * K.CLIGHT) / (K.KBOLTZMANN * K.NM_TO_M * la)
        nl = K.NM_TO_M * la
        lc[n, 2] = (2.0 * K.HPLANCK * K.CLIGHT) / (nl * nl * nl)
        lc[n, 3] = hc_k / la
        lc[n, 4] = twohc / (la * la * la)
        if hmbf_l[0] < la < hmbf_l[-1]:
            lc[n, 5] = spline_eval(hmbf_l, hmbf_a, hmbf_M, la) * 1.0E-21
            lc[n, 6] = 1.0
        if la >= T.HMIN_FF_LAMBDA[-1]:
            lc[n, 7] = -1.0
            lm = la / K.MICRON_TO_NM
            li = 1.0 / lm
            for k in range(1, 6):
                lc[n, 15 + k - 1] = ((lm * lm) * T.HMIN_FF_LONG_A[k] + T.HMIN_FF_LONG_B[k]
                                     + li * (T.HMIN_FF_LONG_C[k] + li * (T.HMIN_FF_LONG_D[k]
                                     + li * (T.HMIN_FF_LONG_E[k] + li * T.HMIN_FF_LONG_F[k]))))
        else:
            lc[n, 7] = frac_index(T.HMIN_FF_LAMBDA, la)
        x = (lam[n] * K.NM_TO_M) / K.CLIGHT
        lc[n, 8] = x * x * x
        xg = ((K.HPLANCK * K.CLIGHT) / (la * K.NM_TO_M)) / (K.E_RYDBERG * (1 * 1))
        lc[n, 9] = math.pow(xg, 0.33333333)
        lc[n, 10] = 2.0 * la * K.NM_TO_M * K.KBOLTZMANN
        lc[n, 11] = (frac_index(T.H2PLUS_FF_LAMBDA, la)
                     if la < T.H2PLUS_FF_LAMBDA[-1] else -1.0)
        lc[n, 12] = (frac_index(T.H2MIN_FF_LAMBDA, la)
                     if la < T.H2MIN_FF_LAMBDA[-1] else -1.0)
        Eev = (K.HPLANCK * K.CLIGHT) / (la * K.NM_TO_M) / K.EV
        lc[n, 13] = (frac_index(T.OH_E, Eev)
                     if not (Eev < T.OH_E[0] or Eev > T.OH_E[-1]) else -1.0)
        lc[n, 14] = (frac_index(T.CH_E, Eev)
                     if not (Eev < T.CH_E[0] or Eev > T.CH_E[-1]) else -1.0)

    # --- bound-free continua: H (Hydrogen_bf) then metals (Metal_bf), in
    #     the order the reference loops over them
    sigma0 = (32.0 / (3.0 * math.sqrt(3.0)) * (K.Q_ELECTRON * K.Q_ELECTRON)
              / (4.0 * K.PI * K.EPSILON_0) / (K.M_ELECTRON * K.CLIGHT)
              * K.HPLANCK / (2.0 * K.E_RYDBERG))
    cont_list = []   # (atom idx, cont, is_H)
    for ia, at in enumerate(atoms):
        if at.active:       # the active atom's continua are in its active set
            continue        # (rh/metal.c:90)
        for c in at.conts:
            cont_list.append((ia, c))
    per_lam = [[] for _ in range(nspect)]   # (global cont index, alpha)
    used = set()
    for gi_, (ia, c) in enumerate(cont_list):
        at = atoms[ia]
        if ia == 0:
            n_eff = math.sqrt(K.E_RYDBERG / (at.E[c.j] - at.E[c.i]))
        for n, la in enumerate(lam):
            if la <= c.lambda0 and la >= c.lam[0]:
                if ia == 0:
                    g_bf = gaunt_bf(la, n_eff, int(at.stage[c.i]) + 1)
                    r = la / c.lambda0
                    a = sigma0 * n_eff * g_bf * (r * r * r)
                elif c.hydrogenic:
                    Z = int(at.stage[c.j])
                    ne_ = Z * math.sqrt(K.E_RYDBERG / (at.E[c.j] - at.E[c.i]))
                    gbf0 = gaunt_bf(c.lambda0, ne_, Z)
                    r = la / c.lambda0
                    a = c.alpha0 * (r * r * r) * gaunt_bf(la, ne_, Z) / gbf0
                else:
                    M = spline_coef(list(c.lam), list(c.alpha))
                    a = spline_eval(list(c.lam), list(c.alpha), M, la)
                per_lam[n].append((gi_, a))
                used.add(gi_)
    used = sorted(used)
    remap = {g: i for i, g in enumerate(used)}
    cont_desc = [cont_list[g] for g in used]
    lam_cont_ptr = [0]
    lam_cont_idx, lam_cont_alpha = [], []
    for n in range(nspect):
        for g, a in per_lam[n]:
            lam_cont_idx.append(remap[g])
            lam_cont_alpha.append(a if cont_list[g][0] == 0 else a * metal_f[n])
        lam_cont_ptr.append(len(lam_cont_idx))
    cont_atom = np.array([ia for ia, _ in cont_desc], np.int32)
    cont_i = np.array([c.i for _, c in cont_desc], np.int32)
    cont_j = np.array([(atoms[0].E.size - 1) if ia == 0 else c.j for ia, c in cont_desc],
                      np.int32)

    # --- passive bound-bound lines of the model atoms would add opacity
    #     (rh/metal.c:174-330); detect and refuse (not implemented)
    for ia, at in enumerate(atoms):
        if at.active:
            continue
        for ln in at.lines:
            dl = ln.lambda0 * ln.qwing * (vchar / K.CLIGHT)
            hit = np.abs(lam - ln.lambda0) <= dl
            if kw.bool("ALLOW_PASSIVE_BB") and hit.any():
                raise NotImplementedError(
                    f"passive bound-bound line of {at.ID} at {ln.lambda0:.3f} nm overlaps "
                    f"the wavelength grid (background atomic lines are not implemented)")
    for m, rng in zip(setup.molecules, setup.mol_ranges):
        if rng is None:
            continue
        lo, hi, qw = rng
        d = lam * qw * (vchar / K.CLIGHT)
        # coarse test in either air or vacuum convention
        if np.any((lam + d >= lo * 0.999) & (lam - d <= hi * 1.001)):
            raise NotImplementedError(
                f"molecular lines of {m.ID} overlap the wavelength grid "
                f"(background molecular lines are not implemented)")

    # --- Kurucz lines contributing at each wavelength (rlk_opacity,
    #     rh/kurucz.c:511-600) and the solver choice (formal.c:78-103)
    rlk = setup.rlk
    nline = len(rlk)
    line_ptr = [0]
    line_idx = []
    solver = np.zeros(nspect, np.int32)
    smode = STOKES_MODES[kw.str("STOKES_MODE")]
    lam0s = np.array([r.lambda0 for r in rlk]) if nline else np.zeros(0)
    model_line_windows = {}
    for ia, at in enumerate(atoms):
        model_line_windows[at.periodic_table] = [
            (ln.lambda0, ln.lambda0 * ln.qwing * (vchar / K.CLIGHT), int(at.stage[ln.i]))
            for ln in at.lines]
    for n, la in enumerate(lam):
        hasline = polar = False
        if nline:
            dlamb_char = la * Q_WING * (vchar / K.CLIGHT)
            if not (la < rlk[0].lambda0 - dlamb_char or la > rlk[-1].lambda0 + dlamb_char):
                for li_, r in enumerate(rlk):
                    if abs(r.lambda0 - la) <= dlamb_char:
                        e = r.pt_index - 1
                        if not (r.stage < el.nstage[e] - 1 and el.abundance_set[e]):
                            continue
                        contributes = True
                        for (l0, dw, st) in model_line_windows.get(e, []):
                            if abs(la - l0) <= dw and st == r.stage:
                                contributes = False
                                break
                        if not contributes:
                            continue
                        line_idx.append(li_)
                        hasline = True
                        if r.polarizable:
                            polar = True
        line_ptr.append(len(line_idx))
        if polar:
            # outside FULL_STOKES a polarizable line still marks the
            # wavelength polarized (pyrh sets atmos.Stokes,
            # pyrh_compute1dray.c:258; kurucz.c:656-659), which keeps it
            # angle-dependent in static columns too (formal.c:98-101)
            solver[n] = 2 if smode == 3 else 3
        elif hasline:
            solver[n] = 1   # refined per column by the 'moving' flag
        else:
            solver[n] = 0

    # --- elements used by lines: partition functions
    line_elems = sorted({r.pt_index - 1 for r in rlk})
    el_map = {e: i for i, e in enumerate(line_elems)}
    npf = el.Tpf.size
    el_nstage = np.array([el.nstage[e] for e in line_elems], np.int32)
    el_pf_ptr, logpf, ionp = [], [], []
    off = 0
    for e in line_elems:
        el_pf_ptr.append(off)
        logpf.append(el.logpf[e].ravel())
        ip = np.zeros(el.nstage[e] * npf)
        ip[:el.nstage[e]] = el.ionpot[e]
        ionp.append(ip)
        off += el.nstage[e] * npf
    el_logpf = np.concatenate(logpf) if logpf else np.zeros(1)
    el_ionpot = np.concatenate(ionp) if ionp else np.zeros(1)

    # --- per-line constants
    hc = K.HPLANCK * K.CLIGHT
    hc_4PI = hc / (4.0 * K.PI)
    la_ = {}
    la_["line_el"] = np.array([el_map[r.pt_index - 1] for r in rlk], np.int32)
    la_["line_stage"] = np.array([r.stage for r in rlk], np.int32)
    la_["line_lambda0"] = lam0s.astype(np.float64)
    la_["line_Ei"] = np.array([r.Ei for r in rlk])
    la_["line_hc_la"] = np.array([(K.HPLANCK * K.CLIGHT) / (r.lambda0 * K.NM_TO_M) for r in rlk])
    la_["line_Bijhc4pi"] = np.array([hc_4PI * r.Bij * r.isotope_frac * r.hyperfine_frac * r.gi
                                     for r in rlk])
    la_["line_twohnu3_c2"] = np.array([r.Aji / r.Bji for r in rlk])
    la_["line_Grad"] = np.array([r.Grad for r in rlk])
    la_["line_GStark"] = np.array([r.GStark for r in rlk])
    la_["line_GvdW"] = np.array([r.GvdWaals for r in rlk])
    la_["line_vdw"] = np.array([r.vdwaals for r in rlk], np.int32)
    la_["line_cross"] = np.array([r.cross for r in rlk])
    la_["line_alpha"] = np.array([r.alpha for r in rlk])
    la_["line_larmor"] = np.array([(K.LARMOR_Q * K.NM_TO_M) * r.lambda0 for r in rlk])
    la_["line_polar"] = np.array([1 if r.polarizable else 0 for r in rlk], np.int32)
    cptr = [0]
    cq, cs, cst = [], [], []
    for r in rlk:
        for (q, s, st) in r.comps:
            cq.append(q)
            cs.append(s)
            cst.append(st)
        cptr.append(len(cq))
    la_["line_comp_ptr"] = np.array(cptr, np.int32)
    la_["comp_q"] = np.array(cq or [0], np.int32)
    la_["comp_shift"] = np.array(cs or [0.0])
    la_["comp_strength"] = np.array(cst or [0.0])

    # --- atom levels
    lvl_ptr = [0]
    lE, lg, ls = [], [], []
    for at in atoms:
        lE.extend(at.E)
        lg.extend(at.g)
        ls.extend(at.stage)
        lvl_ptr.append(len(lE))

    # --- chemical equilibrium nuclei and molecules (rh/chemequil.c:125-200)
    mols = setup.molecules
    nmol_el = [set(m.pt_index) for m in mols]
    nuclei = [e for e in range(len(ELEMENTS)) if any(e in s for s in nmol_el)]
    if nuclei and nuclei[0] != 0:
        raise RHInputError("First nucleus must be H (check H2.molecule)")
    atom_of_el = {at.periodic_table: ia for ia, at in enumerate(atoms)}
    nucl_atom = [atom_of_el.get(e, -1) for e in nuclei]
    for e, ia in zip(nuclei, nucl_atom):
        if ia >= 0 and atoms[ia].stage[0] > 0:
            raise RHInputError(f"Model for element {ELEMENTS[e][0]} does not have a "
                               f"neutral stage needed for molecular formation")
        if ia < 0:
            raise NotImplementedError(
                f"molecule nucleus {ELEMENTS[e][0]} without a model atom "
                f"(Kurucz-pf neutral fraction) is not implemented")
    nucl_abund = [atoms[ia].abundance if ia >= 0 else el.abund[e]
                  for e, ia in zip(nuclei, nucl_atom)]
    mol_eqc = np.zeros((max(len(mols), 1), 8))
    mol_el_nucl = np.zeros((max(len(mols), 1), 4), np.int32)
    mol_el_count = np.zeros((max(len(mols), 1), 4), np.int32)
    for m, mol in enumerate(mols):
        if len(mol.eqc_coef) > 8 or len(mol.pt_index) > 4:
            raise NotImplementedError(f"molecule {mol.ID}: too many coefficients/elements")
        mol_eqc[m, :len(mol.eqc_coef)] = mol.eqc_coef
        for j, (e, c) in enumerate(zip(mol.pt_index, mol.pt_count)):
            mol_el_nucl[m, j] = nuclei.index(e)
            mol_el_count[m, j] = c
    ids = [m.ID for m in mols]

    arrays = dict(
        lamconst=lc, solver=solver,
        lam_line_ptr=np.array(line_ptr, np.int32),
        lam_line_idx=np.array(line_idx or [0], np.int32),
        lam_cont_ptr=np.array(lam_cont_ptr, np.int32),
        lam_cont_idx=np.array(lam_cont_idx or [0], np.int32),
        lam_cont_alpha=np.array(lam_cont_alpha or [0.0]),
        atom_level_ptr=np.array(lvl_ptr, np.int32),
        level_E=np.array(lE), level_g=np.array(lg), level_stage=np.array(ls, np.int32),
        atom_abund=np.array([at.abundance for at in atoms]),
        cont_atom=cont_atom if cont_atom.size else np.zeros(1, np.int32),
        cont_i=cont_i if cont_i.size else np.zeros(1, np.int32),
        cont_j=cont_j if cont_j.size else np.zeros(1, np.int32),
        nucl_atom=np.array(nucl_atom, np.int32), nucl_elem=np.array(nuclei, np.int32),
        nucl_abund=np.array(nucl_abund),
        mol_fit=np.array([m.fit for m in mols], np.int32),
        mol_charge=np.array([m.charge for m in mols], np.int32),
        mol_nnuclei=np.array([m.Nnuclei for m in mols], np.int32),
        mol_ediss=np.array([m.Ediss for m in mols]),
        mol_tmin=np.array([m.Tmin for m in mols]), mol_tmax=np.array([m.Tmax for m in mols]),
        mol_neqc=np.array([len(m.eqc_coef) for m in mols], np.int32),
        mol_eqc=mol_eqc,
        mol_nel=np.array([len(m.pt_index) for m in mols], np.int32),
        mol_el_nucl=mol_el_nucl, mol_el_count=mol_el_count,
        Tpf=el.Tpf.copy(), el_nstage=el_nstage, el_pf_ptr=np.array(el_pf_ptr or [0], np.int32),
        el_logpf=el_logpf, el_ionpot=el_ionpot,
        el_abund=np.array([el.abund[e] for e in line_elems] or [0.0]),
        el_weight=np.array([el.weight[e] for e in line_elems] or [0.0]),
        hminff_theta=np.array(T.HMIN_FF_THETA), hminff_kappa=np.array(T.HMIN_FF_KAPPA),
        h2mff_theta=np.array(T.H2MIN_FF_THETA), h2mff_kappa=np.array(T.H2MIN_FF_KAPPA),
        h2pff_temp=np.array(T.H2PLUS_FF_TEMP), h2pff_kappa=np.array(T.H2PLUS_FF_KAPPA),
        oh_T=np.array(T.OH_T), oh_cross=np.array(T.OH_CROSS),
        ch_T=np.array(T.CH_T), ch_cross=np.array(T.CH_CROSS),
    )
    for k_, v_ in la_.items():
        arrays[k_] = v_ if v_.size else np.zeros(1, v_.dtype)
    scalars = dict(
        stokes_mode=smode, magneto_optical=int(kw.bool("MAGNETO_OPTICAL")),
        vmacro_tresh=vmacro, vmicro_char=vchar,
        wght_per_H=el.wght_per_H, totalAbund=el.totalAbund,
        gravity=K.pow10(4.4) * K.CM_TO_M,
        nspect=nspect, ref_index=ref_index, natom=len(atoms), ncont=len(cont_desc),
        nnuclei=len(nuclei), nmol=len(mols),
        mol_h2=ids.index("H2") if "H2" in ids else -1,
        mol_oh=ids.index("OH") if "OH" in ids else -1,
        mol_ch=ids.index("CH") if "CH" in ids else -1,
        npf=npf, nel=len(line_elems), nline=nline,
    )
    return Model(setup, lam, ref_index, out_index, arrays, scalars, cont_desc)
