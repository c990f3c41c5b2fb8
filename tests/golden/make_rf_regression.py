"""Write tests/golden/rf_adjoint_fixtures.npz: FP64 adjoint response
functions of the library as it stood when their first implementation
(forward-mode dual numbers through every depth step, checked against central
differences of the FP64 synthesis in tests/test_gpu_rf.py) was current.

The reference has no counterpart of these response functions (it only has
d I / d log gf), so this is a regression fixture: a later implementation of
the same derivative (the hand-written reverse sweep) must reproduce it to
rounding.  Runs on the GPU, from the repository root:

    python3 tests/golden/make_rf_regression.py [output.npz]

Stored: three seeded map columns of the 2-line run at mu = 1 and one at
mu = 0.7, every parameter, every third depth and every second wavelength
(float64).
"""
from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
os.environ.setdefault("PYRH_PATH", str(HERE / "rhdata"))

from pyrh_amd import model, workloads as W  # noqa: E402
from pyrh_amd.engine import Engine  # noqa: E402

DEPTH_SEL = slice(0, None, 3)
LAM_SEL = slice(0, None, 2)


def main():
    wave = W.wavelengths_630()
    eng = Engine(model.build_model(model.open_run(str(HERE / "run_lte")), wave))
    out = {"wave": wave}
    for tag, ncol, mu, seed in (("mu1", 3, 1.0, W.SEED_BASE + 71), ("mu07", 1, 0.7, W.SEED_BASE + 72)):
        atm = W.stokes_map(ncol, seed=seed)
        st, rf = eng.rf_lte(atm, mu=mu)
        out[f"{tag}__atmos"] = atm
        out[f"{tag}__mu"] = np.float64(mu)
        out[f"{tag}__stokes"] = st
        out[f"{tag}__rf"] = rf[:, :, DEPTH_SEL, :, LAM_SEL]
        print(tag, rf.shape, float(np.max(np.abs(rf))), flush=True)
    dest = Path(sys.argv[1]) if len(sys.argv) > 1 else HERE / "rf_adjoint_fixtures.npz"
    np.savez_compressed(dest, **out)


if __name__ == "__main__":
    main()
