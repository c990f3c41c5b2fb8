"""Generate tests/golden/mo_fixtures.npz: LTE Stokes with the magneto-optical
terms on (MAGNETO_OPTICAL = TRUE, RH's default, rh/readinput.c:186).

Runs pyrh's rhf1d() from the reference library compiled from its C sources
(oracle/_ref, variant "o2", FP64 4x4 inverse) on tests/golden/run_lte_mo
(run_lte with MAGNETO_OPTICAL = TRUE): the Faraday-Voigt profiles of the
Zeeman components enter the propagation matrix as rho_Q, rho_U, rho_V
(rh/kurucz.c:800-822, 711-715; rh/stokesopac.c:56-86).  The harness in
oracle/ref_harness.c gives the library's magneto-optical background record
room for the four records rh/readj.c:328 copies into it (the library
allocates three); the results read only the first three.  Cases: the
82-point workload column, 8 seeded map columns at mu = 1, and 3 at
mu = 0.7.  The magneto-optical terms change I by up to 20% and Q, U, V by
up to 14% of I_c on these columns.  Needs /root/reference; run from the
repository root:

    PYRH_PATH=tests/golden/rhdata python3 tests/golden/make_golden_mo.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parents[1] / "oracle"))

import refrun  # noqa: E402
from pyrh_amd import workloads as W  # noqa: E402


def cases():
    yield "mo_single", W.single_column()[None], 1.0
    yield "mo_map", W.stokes_map(8, seed=W.SEED_BASE + 61), 1.0
    yield "mo_mu07", W.stokes_map(3, seed=W.SEED_BASE + 62), 0.7


def main():
    if not refrun.ref_available("o2"):
        raise SystemExit("reference build missing: make -C oracle ref")
    wave = W.wavelengths_630()
    out = {"wave": wave}
    for name, atm, mu in cases():
        lam, I, Q, U, V = refrun.run_synth(HERE / "run_lte_mo", atm, wave, mu=mu)[:5]
        st = np.stack([I, Q, U, V], 1)
        off = np.stack(refrun.run_synth(HERE / "run_lte", atm, wave, mu=mu)[1:5], 1)
        out[name + "__atmos"] = atm
        out[name + "__mu"] = np.float64(mu)
        out[name + "__stokes"] = st
        out["lam"] = lam
        Ic = st[:, 0, :1]
        print(name, atm.shape, "MO effect: I", float(np.max(np.abs(st[:, 0] / off[:, 0] - 1))),
              "QUV/Ic", float(np.max(np.abs(st[:, 1:] - off[:, 1:]) / Ic[:, None])), flush=True)
    np.savez_compressed(HERE / "mo_fixtures.npz", **out)


if __name__ == "__main__":
    main()
