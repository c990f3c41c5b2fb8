"""Generate tests/golden/static_fixtures.npz: static columns at a polarizable
line outside FULL_STOKES, and the log gf response functions under ATMOS_ITOP.

Runs pyrh's rhf1d() from the reference library compiled from its C sources
(oracle/_ref, variant "o2").

1. STOKES_MODE = NO_STOKES and FIELD_FREE with VMACRO_TRESH = 0.1 km/s (RH's
   default; tests/golden/run_lte_nostokes_static, run_lte_field_free_static)
   on three seeded map columns: v = 0 (static), |v| = 0.05 km/s (static,
   below the threshold) and v = 1.5 km/s (moving), at mu = 1 and 0.7.  pyrh
   always sets atmos.Stokes (rh/rhf1d/pyrh_compute1dray.c:258), so a
   polarizable Kurucz line marks its wavelengths polarized in the background
   (rh/kurucz.c:656-659, rh/background.c:524-527), which makes them
   angle-dependent (rh/rhf1d/formal.c:98-101): the scalar Bezier solves them
   in the static columns too, where a wavelength with only unpolarizable
   lines would take Feautrier.
2. NO_STOKES with ATMOS_ITOP (tests/golden/run_lte_itop_nostokes) and
   get_atomic_rfs on two log gf parameters, three moving columns at mu = 1
   and 0.7: the inward ray starts at the irradiating intensity
   (rh/rhf1d/bezier_1D.c:378-379) and the emergent ray's log gf recursion
   reads that ray (bezier_1D.c:476-478).

Needs /root/reference; run from the repository root:

    PYRH_PATH=tests/golden/rhdata python3 tests/golden/make_golden_static.py
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

MODES = ("nostokes", "field_free")
# the third wavelength sits in the blue core of Fe I 630.15 nm (air 630.155 nm),
# where the inward ray's intensity weighs on the log gf recursion; the file
# itop.xdr holds one value per wavelength of the merged grid
# {450, 500 (reference), 600, 630.33, 700} nm
ITOP_WAVE = np.array([450.0, 600.0, 630.3292716302619, 700.0])
LOGGF = [(0, -0.71), (1, -0.969)]


def main():
    if not refrun.ref_available("o2"):
        raise SystemExit("reference build missing: make -C oracle ref")
    atm = W.stokes_map(3, seed=W.SEED_BASE + 51)
    atm[0, 3] = 0.0
    atm[1, 3] = 0.05 * np.sin(np.linspace(0.0, 3.0, atm.shape[2]))
    atm[2, 3] = 1.5
    wave = W.wavelengths_630()
    out = {"atmos": atm, "wave": wave}
    for m in MODES:
        for tag, mu in (("mu1", 1.0), ("mu07", 0.7)):
            lam, I, Q, U, V = refrun.run_synth(HERE / f"run_lte_{m}_static", atm, wave, mu=mu)[:5]
            out[f"{m}_{tag}__stokes"] = np.stack([I, Q, U, V], 1)
            out[f"{m}_{tag}__mu"] = np.float64(mu)
            out["lam"] = lam
            # the same columns with the threshold at 0 (every column moving)
            I0 = refrun.run_synth(HERE / f"run_lte_{m}", atm, wave, mu=mu)[1]
            print(m, tag, "static vs moving solver, max rel:",
                  float(np.max(np.abs(I[0] / I0[0] - 1))), flush=True)
    ai = W.stokes_map(3, seed=W.SEED_BASE + 52)
    ai[:, 3] = np.where(np.abs(ai[:, 3]) < 0.2, 0.5, ai[:, 3])
    out["itop_atmos"] = ai
    out["itop_wave"] = ITOP_WAVE
    out["itop_loggf"] = np.array(LOGGF)
    for tag, mu in (("mu1", 1.0), ("mu07", 0.7)):
        lam, I, Q, U, V, rf = refrun.run_synth(HERE / "run_lte_itop_nostokes", ai, ITOP_WAVE,
                                               mu=mu, rf=True, loggf=LOGGF)
        out[f"itop_rf_{tag}__stokes"] = np.stack([I, Q, U, V], 1)
        out[f"itop_rf_{tag}__rf"] = rf
        out[f"itop_rf_{tag}__mu"] = np.float64(mu)
        out["itop_lam"] = lam
        rf0 = refrun.run_synth(HERE / "run_lte_nostokes", ai, ITOP_WAVE, mu=mu, rf=True,
                               loggf=LOGGF)[5]
        print("itop rf", tag, "effect of the irradiation on the RFs, rel to max:",
              float(np.max(np.abs(rf - rf0)) / np.max(np.abs(rf))), flush=True)
    np.savez_compressed(HERE / "static_fixtures.npz", **out)


if __name__ == "__main__":
    main()
