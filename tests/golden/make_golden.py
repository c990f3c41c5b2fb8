"""Generate the golden vectors in tests/golden/lte_fixtures.npz.

Runs pyrh's rhf1d() from the reference library compiled from its C
sources (oracle/_ref, see oracle/Makefile) on fixed inputs and stores the
inputs and the emergent Stokes vectors.  Variant "o2" inverts the 4x4
DELO-Bezier matrix in double precision (the parity target of the FP64
path); "o1" is the library as shipped (float inverse), kept to quantify
that difference.  Needs /root/reference; run from the repository root:

    PYRH_PATH=tests/golden/rhdata python3 tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parents[1] / "oracle"))

import refrun  # noqa: E402
from pyrh_amd import workloads as W  # noqa: E402

import falc  # noqa: E402

RUN = {"lte": HERE / "run_lte", "nostokes": HERE / "run_lte_nostokes",
       "unpol": HERE / "run_lte_unpol", "falc": HERE / "run_falc"}


def cases():
    wave = W.wavelengths_630()
    c0 = W.single_column()
    mp = W.stokes_map(16, seed=W.SEED_BASE + 101)
    yield "single", "lte", c0[None], wave, 1.0, 0
    yield "map16", "lte", mp, wave, 1.0, 0
    yield "mu07", "lte", mp[:4], wave, 0.7, 0
    # column-mass scale: log10 of the reference's own column mass [g cm^-2]
    _, _, _, _, _, cap = refrun.run_synth(RUN["lte"], c0, wave, capture=True)
    cmass_si = [c for c in cap if c[0] == "cmass"][0][5]
    cm = c0.copy()
    cm[0] = np.log10(cmass_si / 10.0)
    yield "cmass", "lte", cm[None], wave, 1.0, 1
    # wavelengths away from the lines: Feautrier branch (formal.c:287-310)
    wave_f = np.concatenate([wave, [627.00, 633.60]])
    yield "feautrier", "lte", mp[:2], wave_f, 1.0, 0
    yield "nostokes", "nostokes", mp[:4], wave, 1.0, 0
    # unpolarizable lines (Armstrong profile); static columns take the
    # Feautrier branch with line opacity, moving ones the scalar Bezier
    st = mp[:2].copy()
    st[:, 3] = 0.0
    yield "unpol_static", "unpol", st, wave, 1.0, 0
    mv = mp[2:4].copy()
    mv[:, 3] = np.where(np.abs(mv[:, 3]) < 0.2, 0.5, mv[:, 3])
    yield "unpol_moving", "unpol", mv, wave, 1.0, 0
    # the reference's benchmark case: FALC, depth-dependent velocities, B = 0,
    # line records without orbital quantum numbers (Unsold broadening)
    fa = falc.atmosphere()[None]
    yield "falc", "falc", fa, falc.wavelengths(), 1.0, 0
    fb = fa.copy()
    fb[0, 5] = 1500.0
    yield "falc_mu05", "falc", fb, falc.wavelengths(), 0.5, 0


# pyrh.compute1d keyword overrides (pyrh.pyx:537-551): log gf and wavelength
# shifts of Kurucz lines by index, element abundances by atomic number
# (the reference carries d I / d log gf only through the scalar Bezier
# solver, bezier_1D.c:416-536: the NO_STOKES case with moving columns has
# them, the FULL_STOKES one returns zeros)
OVERRIDES = {
    "overrides": dict(loggf=[(0, -0.5), (1, -1.2)], lam=[(1, 0.002)], abun=[(26, 7.30)]),
    "rf_nostokes": dict(loggf=[(0, -0.71), (1, -0.969)], lam=[(0, 0.0)]),
}


def override_cases():
    wave = W.wavelengths_630()
    mp = W.stokes_map(16, seed=W.SEED_BASE + 101)
    yield "overrides", "lte", mp[:3], wave, 1.0, 0
    mv = mp[4:7].copy()
    mv[:, 3] = np.where(np.abs(mv[:, 3]) < 0.2, 0.5, mv[:, 3])
    yield "rf_nostokes", "nostokes", mv, wave, 1.0, 0


def main():
    if not refrun.ref_available("o2"):
        raise SystemExit("reference build missing: make -C oracle ref")
    os.environ.setdefault("PYRH_PATH", str(HERE / "rhdata"))
    out = {}
    for name, run, atm, wave, mu, scale in cases():
        lam, I, Q, U, V = refrun.run_synth(RUN[run], atm, wave, mu=mu, atm_scale=scale,
                                           variant="o2")
        out[f"{name}__run"] = np.array(run)
        out[f"{name}__atmos"] = atm
        out[f"{name}__wave"] = wave
        out[f"{name}__mu"] = np.array(mu)
        out[f"{name}__scale"] = np.array(scale)
        out[f"{name}__lam"] = lam
        out[f"{name}__stokes_o2"] = np.stack([I, Q, U, V], axis=1)
        if name == "single":
            _, I1, Q1, U1, V1 = refrun.run_synth(RUN[run], atm, wave, mu=mu,
                                                 atm_scale=scale, variant="o1")
            out[f"{name}__stokes_o1"] = np.stack([I1, Q1, U1, V1], axis=1)
            _, _, _, _, _, cap = refrun.run_synth(RUN[run], atm, wave, mu=mu,
                                                  atm_scale=scale, capture=True)
            out[f"{name}__height"] = [c for c in cap if c[0] == "height"][0][5]
            ref = [c for c in cap if c[0] == "bg_chi" and c[2] == 0 and c[4] == 1][0][5]
            out[f"{name}__chiref"] = ref
        print(name, atm.shape, lam.size)
    for name, run, atm, wave, mu, scale in override_cases():
        ov = OVERRIDES[name]
        lam, I, Q, U, V, rf = refrun.run_synth(RUN[run], atm, wave, mu=mu, atm_scale=scale,
                                               variant="o2", rf=True, **ov)
        out[f"{name}__run"] = np.array(run)
        out[f"{name}__atmos"] = atm
        out[f"{name}__wave"] = wave
        out[f"{name}__mu"] = np.array(mu)
        out[f"{name}__scale"] = np.array(scale)
        out[f"{name}__lam"] = lam
        out[f"{name}__stokes_o2"] = np.stack([I, Q, U, V], axis=1)
        for key in ("loggf", "lam", "abun"):
            pairs = np.array(ov.get(key, []), dtype=np.float64).reshape(-1, 2)
            out[f"{name}__ov_{key}"] = pairs
        # the reference's analytic d I / d log gf (kurucz.c:224-260), per
        # column; its d I / d lambda0 columns are allocated but never filled
        # (uninitialised memory), so only the log gf part is kept
        out[f"{name}__rf_loggf"] = rf[:, :, :len(ov.get("loggf", []))]
        print(name, atm.shape, lam.size, rf.shape)
    np.savez_compressed(HERE / "lte_fixtures.npz", **out)


if __name__ == "__main__":
    main()
