"""Algorithmic work of the LTE Stokes path, frozen.

The unit of work is one emergent Stokes point per (column, wavelength, mu).
Per ray, over depth indices k = 1 .. N-1:

    F = sum_k [ F_FS(solver) + F_CONT
                + sum_{lines in window} ( F_LINE + sum_{components} (F_V(region) + F_SEL) ) ]

plus, once per column, F_HOIST per (line, depth) for the line populations and
damping.  Flops are IEEE add/sub/mul/div (an FMA is 2); exp/sin/cos/pow are
counted apart as transcendentals.  The F_FS figures are structure-aware
minimum counts of one depth step, not the dense 4x4 count of the reference
(rh/rhf1d/bezier_1D.c:169-264, ~1430 with the Cramer inverse), so that a
dense or redundant implementation cannot raise its roofline fraction.  The
Humlicek costs are per region of rh/humlicek.c:28-120; the region of every
component evaluation is replayed by the library's work census
(rhg_count_work) on the same inputs, so F is counted, not estimated.
"""
from __future__ import annotations

import numpy as np

# census slots (include/pyrh_amd.h enum rhg_work)
WORK_SLOTS = ("stokes", "bezier", "feautrier", "line", "voigt_I", "voigt_II",
              "voigt_III", "voigt_IV", "armstrong", "gauss", "coldepth")

F_FS = {"stokes": 384, "bezier": 90, "feautrier": 25}
F_CONT = 40
F_LINE = 25
F_V = {"voigt_I": 20, "voigt_II": 34, "voigt_III": 68, "voigt_IV": 120}
F_SEL = 3
F_ARMSTRONG = 60
F_GAUSS = 4
F_HOIST = 10

# MI355X published peaks
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def census_dict(counts) -> dict:
    c = np.asarray(counts, dtype=np.int64)
    return {name: int(c[i]) for i, name in enumerate(WORK_SLOTS)}


def flops(counts, nline: int) -> dict:
    """FLOPs of one batch by stage, from a census (rhg_count_work)."""
    c = census_dict(counts)
    points = c["stokes"] + c["bezier"] + c["feautrier"]
    formal = sum(F_FS[s] * c[s] for s in F_FS)
    lines = F_LINE * c["line"]
    voigt = sum((F_V[r] + F_SEL) * c[r] for r in F_V) + F_ARMSTRONG * c["armstrong"] \
        + F_GAUSS * c["gauss"]
    cont = F_CONT * points
    hoist = F_HOIST * nline * c["coldepth"]
    return {
        # k_formal: depth recursion + line opacity and profiles of the ray
        "formal_kernel": formal + lines + voigt,
        "continuum_kernel": cont,
        "state_kernel": hoist,
        "total": formal + lines + voigt + cont + hoist,
        "rays_x_depth": points,
        "census": c,
    }


def voigt_transcendentals(counts) -> int:
    c = census_dict(counts)
    return 3 * c["voigt_IV"]
