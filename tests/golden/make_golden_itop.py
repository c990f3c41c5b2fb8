"""Generate the top-irradiation golden vectors in tests/golden/itop_fixtures.npz.

Runs pyrh's rhf1d() from the reference library compiled from its C sources
(oracle/_ref, variant "o2") on tests/golden/run_lte_itop: the LTE run with
ATMOS_ITOP = itop.xdr, the intensity irradiating the top on the merged grid
{450, 500 (reference), 600, 630.2, 700} nm (XDR doubles, one ray,
rh/rhf1d/multiatmos.c:185-222).  The continuum wavelengths take the
Feautrier solver, whose top boundary uses it (rh/rhf1d/feautrier.c:94-96).
Three seeded map columns, one moving, at mu = 1 and mu = 0.7.  Needs
/root/reference; run from the repository root:

    PYRH_PATH=tests/golden/rhdata python3 tests/golden/make_golden_itop.py
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

WAVE = np.array([450.0, 600.0, 630.2, 700.0])


def main():
    atm = W.stokes_map(3, seed=W.SEED_BASE + 50)
    atm[:, 3] = 0.0
    atm[2, 3] = 1.5
    out = {"atmos": atm, "wave": WAVE}
    for tag, mu in (("mu1", 1.0), ("mu07", 0.7)):
        lam, I, Q, U, V = refrun.run_synth(HERE / "run_lte_itop", atm, WAVE, mu=mu)[:5]
        out[tag + "__stokes"] = np.stack([I, Q, U, V], 1)
        out[tag + "__mu"] = np.float64(mu)
        out["lam"] = lam
        lam0, I0 = refrun.run_synth(HERE / "run_lte", atm, WAVE, mu=mu)[:2]
        print(tag, "I with / without irradiation:", np.round(np.asarray(I)[0] / np.asarray(I0)[0], 6),
              flush=True)
    np.savez_compressed(HERE / "itop_fixtures.npz", **out)


if __name__ == "__main__":
    main()
