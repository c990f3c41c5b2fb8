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
