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


# --- k_nlte_bezier (rh/rhf1d/bezier_1D.c:306-541 with the operator,
# rh/fillgamma.c:82-332): FP64 add/sub/mul/div per unit, read off the
# kernel's arithmetic, exp counted apart.  Per ray and depth:
NB_STEP = 98          # the Bezier step: 2 Fritsch-Butland derivatives (16 each),
                      # 4 clamped control points (3 each), dtau (5), the weights of
                      # the dt >= 0.05 branch (37 + exp), I (9), Psi and steps (3)
NB_OPAC = 5           # chi + chi_c, eta + eta_c + sigma J, S = eta / chi
NB_OPAC_LINE = 9      # V = cV phi; chi += V (n_i - g n_j); eta += tw g V n_j
NB_OPAC_CONT = 8      # the same with V = alpha
NB_DEPTH_LINE = 2     # per member and depth (shared by the rays of a block):
NB_DEPTH_CONT = 6     #   w_la, and for continua g_ij (one exp)
NB_EMIT = 3           # J += w_mu I per ray, and dJ once per depth
NB_RATE = 12          # per ray with the operator: Psi/chi, I_eff, w_la mu, Gamma_ji, Gamma_ij
NB_RATE_MEMBER = 12   # per member and ray: eta, chi_cc, U, w_la mu and the two rates
NB_RATE_CROSS = 4     # per coupling term (addtoCoupling)


def nlte_bezier_flops(nm, nd, rb=2):
    """FP64 operations and exps of one pass of k_nlte_bezier over one column
    (every wavelength with an active line, both directions, every ray, nd
    depths), without and with the operator's rate terms.  The set
    composition of each wavelength (members, line or continuum, coupling
    masks) is taken from the atom's active sets as the kernel builds it."""
    nrays = int(nm.muz.size)
    nblocks = -(-nrays // rb)
    out = {"no_op": 0, "op": 0, "exp_no_op": 0, "exp_op": 0}
    for ls in range(nm.as_ptr.size - 1):
        p0, p1 = int(nm.as_ptr[ls]), int(nm.as_ptr[ls + 1])
        if p1 == p0 or not nm.as_bezier[ls]:
            continue
        trs = [int(t) for t in nm.as_tr[p0:p1]]
        lines = [t for t in trs if nm.tr_type[t] == 0]
        conts = [t for t in trs if nm.tr_type[t] != 0]
        m = len(trs)
        opac = nrays * (NB_OPAC + NB_OPAC_LINE * len(lines) + NB_OPAC_CONT * len(conts))
        depth = nblocks * (NB_DEPTH_LINE * len(lines) + NB_DEPTH_CONT * len(conts))
        per_depth = nrays * (NB_STEP + NB_EMIT) + opac + depth
        exps = nrays + nblocks * len(conts)
        # coupling sums: members sharing a lower level, upper level, and
        # upper-lower crossings (set_info's masks)
        ti = [int(nm.tr_i[t]) for t in trs]
        tj = [int(nm.tr_j[t]) for t in trs]
        adds = sum(1 for a in range(m) for b in range(m) if ti[b] == ti[a]) \
            + 2 * sum(1 for a in range(m) for b in range(m) if tj[b] == tj[a]) \
            + sum(1 for a in range(m) for b in range(m) if tj[b] == ti[a])
        cross = sum(1 for a in range(m) for b in range(m) if tj[b] == ti[a])
        rate = nrays * (NB_RATE + NB_RATE_MEMBER * m + adds + NB_RATE_CROSS * cross) \
            + depth + 2 * m
        out["no_op"] += 2 * nd * per_depth
        out["op"] += 2 * nd * (per_depth + rate)
        out["exp_no_op"] += 2 * nd * exps
        out["exp_op"] += 2 * nd * (exps + nblocks * len(conts))
    return out
