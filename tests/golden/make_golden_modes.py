"""Generate the STOKES_MODE golden vectors in tests/golden/modes_fixtures.npz.

Runs pyrh's rhf1d() from the reference library compiled from its C sources
(oracle/_ref, variant "o2") on the LTE run directory with STOKES_MODE =
FIELD_FREE and = POLARIZATION_FREE (tests/golden/run_lte_field_free,
run_lte_polarization_free; otherwise as run_lte_nostokes) on three seeded
map columns, two of them moving, at mu = 1 and mu = 0.7.  With no active
atom both modes solve Stokes I with the scalar solvers on the Zeeman-split
line opacity (rh/rhf1d/formal.c:94-95; adjustStokesMode, rh/zeeman.c:315-317,
leaves the mode alone).  Needs /root/reference; run from the repository
root:

    PYRH_PATH=tests/golden/rhdata python3 tests/golden/make_golden_modes.py
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

MODES = ("field_free", "polarization_free")


def main():
    atm = W.stokes_map(3, seed=W.SEED_BASE + 49)
    atm[:, 3] += np.array([0.0, 1.5, -2.5])[:, None]
    wave = W.wavelengths_630()
    out = {"atmos": atm, "wave": wave}
    for m in MODES:
        for tag, mu in (("mu1", 1.0), ("mu07", 0.7)):
            lam, I, Q, U, V = refrun.run_synth(HERE / f"run_lte_{m}", atm, wave, mu=mu)[:5]
            out[f"{m}_{tag}__stokes"] = np.stack([I, Q, U, V], 1)
            out[f"{m}_{tag}__mu"] = np.float64(mu)
            out["lam"] = lam
            print(m, tag, flush=True)
    np.savez_compressed(HERE / "modes_fixtures.npz", **out)


if __name__ == "__main__":
    main()
