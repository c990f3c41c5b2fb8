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


def hse_flops(hm, layer_iters, newton_per_iter, chem_per_iter):
    """Algorithmic FP64 operations of pyrh's hse() (rh/rhf1d/pyrh_hse.c:67-399)
    for `layer_iters` pressure iterations, counting each add, multiply,
    divide, exp, pow, log and sqrt of the reference's arithmetic as one:

    * a Newton step of get_ne (rh/rhf1d/pyrh_background.c:509-538) over the
      element table: 11 per element (its ne factor, first partition
      function, normalisation of stage 0) + 24 per further stage (partition
      function, Saha factor, derivative, sums, normalisation, error and
      derivative terms) + 20;
    * an iteration of the chemical equilibrium (rh/chemequil.c:262-330): LU
      solution of the N = nuclei + molecules system with one refinement,
      2/3 N^3 + 4 N^2, and 10 per molecule to build it;
    * per pressure iteration: LTE populations of every atom, 9 per level
      (rh/ltepops.c:87-110), 15 per molecular equilibrium constant, 150 for
      the 500 nm opacity and the pressure step.

    newton_per_iter and chem_per_iter are the mean counts per pressure
    iteration (from the oracle's census of a sample)."""
    m = hm.model.scalars
    nel = int(hm.nel)
    nst = int(np.sum(hm.el_nstage)) - nel
    newton = 11 * nel + 24 * nst + 20
    N = int(m["nnuclei"]) + int(m["nmol"])
    chem = 2.0 / 3.0 * N ** 3 + 4 * N * N + 10 * int(m["nmol"])
    nlev = int(hm.model.arrays["atom_level_ptr"][-1])
    layer = 9 * nlev + 15 * int(m["nmol"]) + 150
    per_iter = newton_per_iter * newton + chem_per_iter * chem + layer
    return float(layer_iters) * per_iter, {"newton_step": newton, "chem_iteration": chem,
                                           "layer_fixed": layer, "per_layer_iteration": per_iter}
