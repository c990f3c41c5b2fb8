"""Synthetic workloads (seeded) for the LTE Stokes synthesis.

* ``single_column()``: 82 depths uniform in log tau500 from -6 to +1, a
  smooth semi-empirical-like temperature (about 9000 K at the top, a 4400 K
  minimum near log tau = -3, 6400 K at log tau = 0 and 9300 K at +1, PCHIP
  through those nodes), gas pressure from hydrostatic equilibrium with
  p_g(top) = 0.1 Pa and n_e from LTE ionisation (stored in
  data/lte_atmos82.npz, produced by pyrh's own hse()), the Fe I
  630.15/630.25 nm pair on 112 wavelengths at 2.15 pm air spacing from
  630.089 nm, B = 1000 G, gamma = 45 deg, phi = 30 deg, v_LOS = 0,
  v_mic = 1 km/s, mu = 1.
* ``stokes_map(ncol, seed)``: columns of the same shape with T = T0 plus a
  natural cubic spline through 5 log-tau nodes drawn N(0, 150 K),
  B ~ U(0, 3000 G), cos(gamma) ~ U(-1, 1), phi ~ U(0, pi),
  v_LOS ~ N(0, 2 km/s), v_mic ~ U(0.5, 2 km/s).  n_e and n_H are those of
  the single column (the hydrostatic structure is not re-solved per
  column).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from . import constants as K

DATA = Path(__file__).resolve().parent / "data"
NDEP = 82
LOGTAU = np.linspace(-6.0, 1.0, NDEP)
T_NODES = ([-6.0, -3.0, 0.0, 1.0], [9000.0, 4400.0, 6400.0, 9300.0])
PERTURB_NODES = np.array([-6.0, -4.25, -2.5, -0.75, 1.0])
SEED_BASE = 20260


def t0_profile(logtau=LOGTAU):
    from scipy.interpolate import PchipInterpolator
    return PchipInterpolator(*T_NODES)(logtau)


def wavelengths_630():
    """112 vacuum wavelengths [nm]: air 630.089 + 0.00215 i nm."""
    lam_air = 630.089 + 0.00215 * np.arange(112)
    return np.array(K.air_to_vacuum(lam_air))


def _hse_structure():
    d = np.load(DATA / "lte_atmos82.npz")
    return d["ne_cm3"], d["nH_cm3"]


def single_column():
    ne, nH = _hse_structure()
    atm = np.zeros((9, NDEP))
    atm[0] = LOGTAU
    atm[1] = t0_profile()
    atm[2] = ne
    atm[3] = 0.0
    atm[4] = 1.0
    atm[5] = 1000.0
    atm[6] = np.pi / 4
    atm[7] = np.pi / 6
    atm[8] = nH
    return atm


def stokes_map(ncol, seed=SEED_BASE + 1):
    from scipy.interpolate import CubicSpline
    rng = np.random.Generator(np.random.PCG64(seed))
    ne, nH = _hse_structure()
    T0 = t0_profile()
    atm = np.zeros((ncol, 9, NDEP))
    dT = rng.normal(0.0, 150.0, size=(ncol, PERTURB_NODES.size))
    B = rng.uniform(0.0, 3000.0, ncol)
    cosg = rng.uniform(-1.0, 1.0, ncol)
    phi = rng.uniform(0.0, np.pi, ncol)
    vlos = rng.normal(0.0, 2.0, ncol)
    vmic = rng.uniform(0.5, 2.0, ncol)
    spl = CubicSpline(PERTURB_NODES, dT.T, bc_type="natural")(LOGTAU).T
    atm[:, 0] = LOGTAU
    atm[:, 1] = T0[None, :] + spl
    atm[:, 2] = ne[None, :]
    atm[:, 3] = vlos[:, None]
    atm[:, 4] = vmic[:, None]
    atm[:, 5] = B[:, None]
    atm[:, 6] = np.arccos(cosg)[:, None]
    atm[:, 7] = phi[:, None]
    atm[:, 8] = nH[None, :]
    return atm


def rf_batch(ncol=1024, seed=SEED_BASE + 2):
    """The response-function batch: columns drawn as in the Stokes map,
    from their own seed."""
    return stokes_map(ncol, seed=seed)


# ---------------------------------------------------------------------------
# Mixed-size batch: 400 synthetic Fe I Zeeman lines on a 2000-point grid,
# columns with 64 to 256 depth points

MIXED_NDEP = (64, 96, 128, 192, 256)
MIXED_MU = (1.0, 0.7, 0.4)
MIXED_NLINE = 400
# J of the lower and upper level of every synthetic line (fixed)
MIXED_J = (1.0, 2.0)
_CM1_PER_EV = 8065.543937


def mixed_wavelengths():
    """2000 vacuum wavelengths [nm]: air 620 + 0.01 i nm, i = 0..1999."""
    return np.array(K.air_to_vacuum(620.0 + 0.01 * np.arange(2000)))


def mixed_line_records(nline=MIXED_NLINE, seed=SEED_BASE + 7):
    """Kurucz 160-column records of synthetic Fe I lines: log gf ~ U(-4, 0),
    E_low ~ U(0, 5) eV, lambda0 (air) ~ U(620, 640) nm, Lande factors of
    both levels ~ U(0, 2.5) in the Lande columns (used with LS_LANDE =
    FALSE), J = 1 -> 2, l-numbers 0 -> 1, the broadening constants of the
    Fe I 630.15 nm record."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lam = rng.uniform(620.0, 640.0, nline)
    gf = rng.uniform(-4.0, 0.0, nline)
    e_lo = rng.uniform(0.0, 5.0, nline) * _CM1_PER_EV
    g_lo = rng.uniform(0.0, 2.5, nline)
    g_up = rng.uniform(0.0, 2.5, nline)
    recs = []
    for i in range(nline):
        e_up = e_lo[i] + 1.0e7 / lam[i]
        r = [" "] * 163

        def put(pos, text):
            r[pos:pos + len(text)] = list(text)
        put(0, f"{lam[i]:10.4f}")
        put(11, f"{gf[i]:6.3f}")
        put(18, "26.00")
        put(24, f"{e_lo[i]:11.3f}")
        put(36, f"{MIXED_J[0]:4.1f}")
        put(41, "4s6D5s e5D")
        put(52, f"{e_up:11.3f}")
        put(64, f"{MIXED_J[1]:4.1f}")
        put(69, "5Dsp3P z5P")
        put(79, "  8.08 -5.42 -7.54")
        put(97, "K17  0 0 ")
        put(106, " 0 0.000  0 0.000")
        put(143, f"{int(round(g_lo[i] * 1000)):5d}{int(round(g_up[i] * 1000)):5d}")
        put(153, "     0 0 1")
        recs.append("".join(r))
    return recs


def mixed_run_dir(path, base_run):
    """Write the mixed-batch run directory into `path`: the keyword, atoms
    and molecules files of `base_run` (the LTE run directory), LS_LANDE =
    FALSE, ALLOW_PASSIVE_BB = FALSE, and a Kurucz list with the synthetic
    lines.  Returns path."""
    import os
    os.makedirs(path, exist_ok=True)
    kw = open(os.path.join(base_run, "keyword.input")).read().splitlines()
    out = []
    for ln in kw:
        key = ln.split("=")[0].strip() if "=" in ln else ""
        if key in ("KURUCZ_DATA", "LS_LANDE", "ALLOW_PASSIVE_BB"):
            continue
        out.append(ln)
    # the model atoms' own passive lines (Si I 635.7 nm ...) stay out, as
    # ALLOW_PASSIVE_BB = FALSE does in the reference
    out += ["  KURUCZ_DATA = kurucz.input", "  LS_LANDE = FALSE", "  ALLOW_PASSIVE_BB = FALSE"]
    with open(os.path.join(path, "keyword.input"), "w") as f:
        f.write("\n".join(out) + "\n")
    for name in ("atoms.input", "molecules.input"):
        with open(os.path.join(base_run, name)) as src, open(os.path.join(path, name), "w") as dst:
            dst.write(src.read())
    # the list names the line file relative to the run directory, which
    # the reference opens from its working directory
    with open(os.path.join(path, "fe_synthetic400"), "w") as f:
        f.write("\n".join(mixed_line_records()) + "\n")
    with open(os.path.join(path, "kurucz.input"), "w") as f:
        f.write("fe_synthetic400\n")
    return path


def mixed_batch(ncol, seed=SEED_BASE + 8, ndep_choices=MIXED_NDEP):
    """Columns [ncol, 9, max(ndep_choices)] and their depth counts: each
    column's depth count drawn uniformly from ndep_choices, its points
    uniform in log tau500 from -6 to +1, T = T0 plus the map's perturbation,
    n_e and n_H of the hydrostatic structure interpolated in log, B, angles
    and velocities drawn as in the map; zero past the column's depth."""
    from scipy.interpolate import CubicSpline
    rng = np.random.Generator(np.random.PCG64(seed))
    ne82, nH82 = _hse_structure()
    nmax = max(ndep_choices)
    nd = rng.choice(np.array(ndep_choices), size=ncol)
    dT = rng.normal(0.0, 150.0, size=(ncol, PERTURB_NODES.size))
    B = rng.uniform(0.0, 3000.0, ncol)
    cosg = rng.uniform(-1.0, 1.0, ncol)
    phi = rng.uniform(0.0, np.pi, ncol)
    vlos = rng.normal(0.0, 2.0, ncol)
    vmic = rng.uniform(0.5, 2.0, ncol)
    atm = np.zeros((ncol, 9, nmax))
    grids = {}
    for n in ndep_choices:
        lt = np.linspace(-6.0, 1.0, n)
        grids[n] = (lt, t0_profile(lt), np.exp(np.interp(lt, LOGTAU, np.log(ne82))),
                    np.exp(np.interp(lt, LOGTAU, np.log(nH82))))
    for c in range(ncol):
        n = int(nd[c])
        lt, T0, ne, nH = grids[n]
        atm[c, 0, :n] = lt
        atm[c, 1, :n] = T0 + CubicSpline(PERTURB_NODES, dT[c], bc_type="natural")(lt)
        atm[c, 2, :n] = ne
        atm[c, 3, :n] = vlos[c]
        atm[c, 4, :n] = vmic[c]
        atm[c, 5, :n] = B[c]
        atm[c, 6, :n] = np.arccos(cosg[c])
        atm[c, 7, :n] = phi[c]
        atm[c, 8, :n] = nH
    return atm, nd.astype(np.int32)
