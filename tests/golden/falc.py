"""The FALC atmosphere and wavelength grid of the reference's benchmark
(benchmark/synth.py:11-37 and :73-77): a SPINOR-format model (falc.dat, 57
depth points) cast to pyrh's 9-row layout, and 301 air wavelengths around
630.2 nm converted to vacuum."""
from pathlib import Path

import numpy as np

RUN = Path(__file__).resolve().parent / "run_falc"
K_B = 1.380649e-23          # scipy.constants.k, as the benchmark script uses


def atmosphere():
    a = np.loadtxt(RUN / "falc.dat", skiprows=1).T
    atm = np.empty((9, a.shape[1]))
    atm[0] = a[0]                                   # log tau500
    atm[1] = a[2]                                   # T [K]
    atm[2] = a[4] / 10 / K_B / a[2] / 1e6           # ne [cm^-3] from Pe [dyn cm^-2]
    atm[3] = a[9] / 1e5                             # vLOS [km/s]
    atm[4] = a[8] / 1e5                             # vmic [km/s]
    atm[5] = a[7]                                   # B [G]
    atm[6] = a[-2]                                  # gamma [rad]
    atm[7] = a[-1]                                  # chi [rad]
    atm[8] = (a[3] - a[4]) / 10 / K_B / a[2] / 1e6 / 1.26   # nH [cm^-3]
    return atm


def wavelengths():
    air = np.linspace(630.2 - 0.15, 630.2 + 0.15, num=301)
    sig = (1.0e7 / air) ** 2
    return air * (1.0000834213 + 2.406030e6 / (1.3e10 - sig) + 1.5997e4 / (3.89e9 - sig))


def stored_spectrum():
    """benchmark/hinode_falc_pyrh.spec: I, Q, U, V, lambda (7 digits)."""
    return np.loadtxt(RUN / "hinode_falc_pyrh.spec").T
