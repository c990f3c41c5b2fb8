"""Write the mixed-batch run directory tests/golden/run_mixed and its golden
vectors tests/golden/mixed_fixtures.npz.

The run directory holds the LTE run's keyword/atoms/molecules files with
LS_LANDE = FALSE and ALLOW_PASSIVE_BB = FALSE, and 400 synthetic Fe I lines
with Lande factors in the Kurucz records (pyrh_amd.workloads.mixed_*).  The
reference's rhf1d() compiled from its C sources (oracle/_ref, variant "o2",
FP64 4x4 inverse) synthesises one column of each depth count (64, 96, 128,
192, 256 points) on the 2000-point grid at mu = 1 and mu = 0.4; the columns
of one depth count go to the reference together.  Needs /root/reference;
run from the repository root:

    PYRH_PATH=tests/golden/rhdata python3 tests/golden/make_golden_mixed.py
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

RUN = HERE / "run_mixed"


def main():
    W.mixed_run_dir(str(RUN), str(HERE / "run_lte"))
    wave = W.mixed_wavelengths()
    atm, nd = W.mixed_batch(64, seed=W.SEED_BASE + 108)
    pick = [int(np.flatnonzero(nd == n)[0]) for n in W.MIXED_NDEP]
    atm, nd = atm[pick], nd[pick]
    out = {"wave": wave}
    for name, mu in (("mixed_mu1", 1.0), ("mixed_mu04", 0.4)):
        st = None
        for c in range(len(pick)):
            n = int(nd[c])
            lam, I, Q, U, V = refrun.run_synth(RUN, atm[c:c + 1, :, :n], wave, mu=mu,
                                               variant="o2", timeout=3600)
            s = np.stack([I, Q, U, V], axis=1)
            st = s if st is None else np.concatenate([st, s])
        out[f"{name}__atmos"] = atm
        out[f"{name}__ndep"] = nd
        out[f"{name}__mu"] = mu
        out[f"{name}__stokes_o2"] = st
        out[f"{name}__lam"] = lam
        print(name, st.shape, flush=True)
    np.savez_compressed(HERE / "mixed_fixtures.npz", **out)


if __name__ == "__main__":
    main()
