"""Generate the NLTE golden vectors in tests/golden/nlte_fixtures.npz.

Runs pyrh's rhf1d() from the reference library compiled from its C sources
(oracle/_ref, variant "o2") on the NLTE run directory tests/golden/run_nlte
(the Ca II-like five-level atom with continuum, active, accelerated lambda
iteration to a maximum relative population change of 1e-3) and stores the
inputs, the converged populations of the active atom, the iteration count
and last change, the merged wavelength grid, the iteration history and the
emergent spectrum that rhf1d() returns afterwards at the case's mu.
A second set, tests/golden/nlte_mg_fixtures.npz, solves the neutral Mg I
atom of rh/Atoms/MgI_6level.atom (run directory tests/golden/run_nlte_mg):
its Mg b lines take Barklem van der Waals broadening from the ABO tables
(rh/barklem.c:202-313, rh/broad.c:125-136), which the ionised Ca II lines
cannot (the reference falls back to UNSOLD for them).
Needs /root/reference; run from the repository root:

    PYRH_PATH=tests/golden/rhdata python3 tests/golden/make_golden_nlte.py [caii] [mg]
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

RUN = HERE / "run_nlte"
WAVE = np.array([393.4776, 396.9591, 854.44])
RUN_MG = HERE / "run_nlte_mg"
WAVE_MG = np.array([285.3, 517.3, 518.5])


def cases():
    mp = W.stokes_map(6, seed=W.SEED_BASE + 3)
    yield "map6", mp, 1.0
    # the single workload column and a slow, strongly perturbed one
    c0 = W.single_column()[None]
    hot = mp[:1].copy()
    hot[:, 1] *= 1.08
    yield "single", c0, 1.0
    yield "hot", hot, 1.0
    # an inclined final ray
    yield "mu07", mp[2:4].copy(), 0.7


def cases_mg():
    mp = W.stokes_map(3, seed=W.SEED_BASE + 7)
    yield "map3", mp, 1.0
    yield "mu07", mp[1:2].copy(), 0.7


def solve(run, wave, case_list, fname):
    out = {}
    for name, atm, mu in case_list:
        res = refrun.run_synth(run, atm, wave, mu=mu, atm_scale=0, variant="o2",
                               capture=True, timeout=3600)
        cap = res[-1]
        ncol = atm.shape[0]
        # records: pops_n with (column, atom, level) in (col, nspect, mu)
        nl = len([r for r in cap if r[0] == "pops_n" and r[1] == 0])
        pops = np.zeros((ncol, nl, atm.shape[2]))
        nstar = np.zeros((ncol, nl, atm.shape[2]))
        niter = np.zeros(ncol, np.int32)
        dpops = np.zeros(ncol)
        hist = np.zeros((ncol, 100))
        for c in range(ncol):
            for r in cap:
                if r[0] == "pops_n" and r[1] == c and r[2] == 0:
                    pops[c, r[3]] = r[5]
                if r[0] == "pops_nstar" and r[1] == c and r[2] == 0:
                    nstar[c, r[3]] = r[5]
            its = [r[5] for r in cap if r[0] == "iter" and r[1] == c]
            niter[c] = len(its)
            dpops[c] = its[-1][1]
            hist[c, :len(its)] = [x[1] for x in its]
        lam = [r[5] for r in cap if r[0] == "lambda"][0]
        # each line's span on the merged grid as SortLambda left it
        spans = {}
        for r in cap:
            if r[0] == "line_lambda" and r[1] == 0:
                spans.setdefault(r[2], (int(np.searchsorted(lam, r[5][0])), r[5].size))
        out[f"{name}__line_span"] = np.array([spans[k] for k in sorted(spans)], np.int32)
        out[f"{name}__atmos"] = atm
        out[f"{name}__pops"] = pops
        out[f"{name}__nstar"] = nstar
        out[f"{name}__niter"] = niter
        out[f"{name}__dpops"] = dpops
        out[f"{name}__history"] = hist
        out[f"{name}__lam"] = lam
        out[f"{name}__mu"] = mu
        out[f"{name}__stokes"] = np.stack(res[1:5], axis=1)     # [ncol, 4, nlw]
        out[f"{name}__lamout"] = res[0]
        print(name, atm.shape, "iterations", niter.tolist())
    out["wave"] = wave
    np.savez_compressed(HERE / fname, **out)


def main():
    which = sys.argv[1:] or ["caii", "mg"]
    if "caii" in which:
        solve(RUN, WAVE, cases(), "nlte_fixtures.npz")
    if "mg" in which:
        solve(RUN_MG, WAVE_MG, cases_mg(), "nlte_mg_fixtures.npz")


if __name__ == "__main__":
    main()
