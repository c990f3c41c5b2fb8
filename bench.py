"""Throughput of the LTE Stokes synthesis on one MI355X.

Workload: the seeded 128x128 = 16384-column LTE Stokes map (82 depth points
uniform in log tau500 from -6 to +1, the Fe I 630.15/630.25 nm pair with
Zeeman splitting, 112 wavelengths from 630.089 nm air at 2.15 pm, mu = 1),
synthetic atmospheres from pyrh_amd.workloads.  One step = one synthesis of
the whole map with the atmosphere already resident in HBM.

Prints one JSON line: columns/s of the FP32 form (the plain FP64 form beside
it), ms per step, the roofline of the dominant kernel (k_formal_stokes:
algorithmic FLOPs from the library's work census over its HIP-event launch
time on the library's stream), and a CPU baseline timed on this host on a
bounded sample of the same columns.  The response-function leg (object
response_functions) times the adjoint of a 1024-column batch in the FP32
form and in FP64 against central differences of the plain FP64 synthesis
(2 x 6 x 82 perturbed syntheses per column) and reports the errors.  The
NLTE leg (object nlte) solves the populations of a Ca II-like atom in 4096
map columns to convergence and reports columns/s, the iteration histogram
and the roofline of its formal-solution kernel.  The mixed leg (object
mixed) synthesises 8192 columns of 64-256 depth points with 400 Zeeman
lines on 2000 wavelengths at three mu, with adjoint response functions on
a subset and the oracle on 128 columns.  The HSE leg (object hse) builds
the hydrostatic structure (pyrh's hse()) of 16384 map temperature columns
and reports columns/s, the pressure-iteration census and the roofline of
its kernel.

    python bench.py [--steps K] [--warmup W] [--ncol N] [--no-cpu] [--no-rf]
                    [--no-nlte] [--nlte-ncol N] [--nlte-only]
                    [--no-mixed] [--mixed-ncol N] [--mixed-only]
                    [--no-hse] [--hse-ncol N] [--hse-only]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
os.environ.setdefault("PYRH_PATH", str(ROOT / "tests" / "golden" / "rhdata"))

RUN_DIR = ROOT / "tests" / "golden" / "run_lte"
PROFILES = ROOT / "profiles"
METRIC = ("columns/s Stokes+RF, FP32 adjoint vs plain FP64 GPU at ≤1e-4 rel err; CPU ref")
WORKLOAD = "lte_stokes_map_16384"
RF_WORKLOAD = "lte_stokes_rf_1024"
# central-difference steps of the finite-difference response functions, in
# the atmosphere's units (K, km/s, G, rad, rad, km/s)
RF_STEP = {"T": 1.0, "vlos": 0.01, "B": 1.0, "gamma": 0.01, "chi": 0.01, "vmic": 0.01}
# NLTE leg: the Ca II-like five-level atom with continuum (rh/Atoms/CaII.atom
# with its lines in complete redistribution, 101 points each), ACTIVE, in the
# run directory tests/golden/run_nlte (5 Gauss-Legendre rays, Ng 2/10/3,
# ITER_LIMIT 1e-3), on 4096 map columns
NLTE_RUN = ROOT / "tests" / "golden" / "run_nlte"
NLTE_WORKLOAD = "nlte_caii_4096"
NLTE_WAVE = (393.4776, 396.9591, 854.44)


def _traffic_from_profile(ncol):
    """HBM bytes per k_formal launch from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    f = PROFILES / "k_formal_lte_stokes_map_16384_hbm.md"
    if not f.exists() or ncol != 16384:
        return None
    m = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", f.read_text())
    return float(m.group(1)) if m else None


def _profile_kernels():
    """One launch of each kernel of the bench at its workload's shape, inputs
    in HBM, then one 1 GiB device-to-device copy (the counter calibration):
    the FP32-form and plain FP64 syntheses of the 16384-column map, the
    FP32-form response functions of the 1024-column batch (all six
    parameters), an NLTE solve of 512 columns of the Ca II-like atom, and
    hse() on 16384 columns."""
    import torch
    from pyrh_amd import hse as HS, model as M, nlte as NL, workloads as W
    from pyrh_amd.engine import Engine
    dev = torch.device("cuda:0")
    setup = M.open_run(str(RUN_DIR))
    wave = W.wavelengths_630()
    eng = Engine(M.build_model(setup, wave))
    ncol = 16384
    atm = torch.from_numpy(W.stokes_map(ncol, seed=W.SEED_BASE + 1)).to(dev)
    nd = atm.shape[2]
    ndep = torch.full((ncol,), nd, dtype=torch.int32, device=dev)
    out = torch.empty((ncol, 4, eng.nout), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    for prec in (1, 0):
        eng.synth_lte_device(atm.data_ptr(), ndep.data_ptr(), ncol, nd, out.data_ptr(),
                             precision=prec)
    print(f"profile: synthesis FP32 and FP64, {ncol} columns", flush=True)
    nrf = 1024
    rf = torch.empty((6, nd, 4, eng.nout, nrf), dtype=torch.float32, device=dev)
    eng.rf_lte_device(atm.data_ptr(), ndep.data_ptr(), nrf, nd, out.data_ptr(), rf.data_ptr(),
                      precision=1)
    print(f"profile: response functions FP32, {nrf} columns", flush=True)
    del rf
    eng.close()
    nm = NL.build_nlte(M.open_run(str(NLTE_RUN)), np.array(NLTE_WAVE))
    en = Engine(nm.lte)
    nn = 512
    natm = torch.from_numpy(W.stokes_map(nn, seed=W.SEED_BASE + 5)).to(dev)
    nnd = torch.full((nn,), nd, dtype=torch.int32, device=dev)
    pops = torch.empty((nn, nm.nlevel, nd), dtype=torch.float64, device=dev)
    nit = torch.empty(nn, dtype=torch.int32, device=dev)
    dp = torch.empty(nn, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    en.nlte_solve_device(nm, natm.data_ptr(), nnd.data_ptr(), nn, nd, pops.data_ptr(),
                         nit.data_ptr(), dp.data_ptr())
    print(f"profile: NLTE, {nn} columns", flush=True)
    en.close()
    hm = HS.build_hse(M.open_run(str(RUN_DIR)))
    eh = Engine(hm.model)
    T = torch.from_numpy(np.ascontiguousarray(W.stokes_map(ncol, seed=W.SEED_BASE + 44)[:, 1]))
    sc = torch.from_numpy(np.ascontiguousarray(np.broadcast_to(W.LOGTAU, tuple(T.shape))))
    T, sc = T.to(dev), sc.to(dev)
    pg = torch.full((ncol,), 0.1, dtype=torch.float64, device=dev)
    hs = [torch.empty((ncol, nd), dtype=torch.float64, device=dev) for _ in range(4)]
    it = torch.empty((ncol, nd), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    eh.hse_device(hm, sc.data_ptr(), T.data_ptr(), pg.data_ptr(), ndep.data_ptr(), ncol, nd,
                  *(h.data_ptr() for h in hs), 0, it.data_ptr())
    print(f"profile: hse, {ncol} columns", flush=True)
    eh.close()
    x = torch.empty(1 << 30, dtype=torch.uint8, device=dev)
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    print("profile: 1 GiB copy done", flush=True)


def _host_cpu():
    """The host CPU model and how many logical CPUs this process may use;
    the baselines use at most 16 of them, the host share of one GPU."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}; {len(os.sched_getaffinity(0))} logical CPUs in the affinity mask"


def _cpu_baseline(setup, wave, atm, nworkers):
    """The reference CPU path (rhf1d built from the reference sources, the
    shipped float-inverse variant) on a bounded sample, nworkers processes
    at once; the oracle restatement with OpenMP if that build is absent."""
    sys.path.insert(0, str(ROOT / "oracle"))
    per_worker = 32
    ns = min(atm.shape[0], per_worker * nworkers)
    sample = atm[:ns]
    import refrun
    if refrun.ref_available("o1"):
        from concurrent.futures import ThreadPoolExecutor
        chunks = np.array_split(sample, nworkers)
        t0 = time.perf_counter()
        with ThreadPoolExecutor(nworkers) as ex:
            list(ex.map(lambda a: refrun.run_synth(RUN_DIR, a, wave, variant="o1"), chunks))
        dt = time.perf_counter() - t0
        return {"value": ns / dt, "unit": "columns/s", "cores": nworkers, "cpu": _host_cpu(), "kind": "reference",
                "sample": f"{ns} columns of the map, rhf1d() per column (input files re-read "
                          f"per call as pyrh does), {nworkers} processes, {dt:.1f} s wall"}
    import pyoracle
    from pyrh_amd import model as M
    m = M.build_model(setup, wave)
    t0 = time.perf_counter()
    pyoracle.synth_lte(m, sample, nthreads=nworkers)
    dt = time.perf_counter() - t0
    return {"value": ns / dt, "unit": "columns/s", "cores": nworkers, "cpu": _host_cpu(), "kind": "port",
            "sample": f"{ns} columns of the map, C restatement, OpenMP {nworkers} threads, "
                      f"{dt:.1f} s wall"}


def _err_stats(x, ref, thr, names=None):
    """Error of x against ref per parameter, relative to each (parameter,
    depth, column)'s largest |ref| over Stokes and wavelength; arrays in the
    library's layout [param, depth, 4, nout, ncol].  Median and 99th
    percentile (from a strided sample) and the exact share above thr."""
    import torch
    from pyrh_amd.engine import RF_PARAMS
    out = {}
    for ip, name in enumerate(names or RF_PARAMS):
        r = ref[ip]
        scale = r.abs().amax(dim=(1, 2), keepdim=True)
        e = (x[ip].double() - r).abs() / scale.clamp_min(1e-300)
        e = e.masked_select((scale > 0).expand_as(e))
        smp = e[:: max(1, e.numel() // 8_000_000)]
        q = torch.quantile(smp, torch.tensor([0.5, 0.99], dtype=smp.dtype, device=smp.device))
        out[name] = {"p50": float(f"{q[0].item():.3g}"), "p99": float(f"{q[1].item():.3g}"),
                     f"share_gt_{thr:g}": float(f"{(e > thr).double().mean().item():.3g}")}
    return out


def _rf_leg(eng, W, torch, dev, reps):
    """Response functions of the 1024-column batch: the adjoint (FP32 form
    and FP64) against central differences of the plain FP64 synthesis on the
    GPU, 2 x 6 x ndep perturbed syntheses per column, batched 16 depths x
    1024 columns per launch."""
    from pyrh_amd.engine import RF_PARAMS, RF_ROWS, RF_NPARAM
    atm = W.rf_batch()
    ncol, nd = atm.shape[0], atm.shape[2]
    nout = eng.nout
    d_atm = torch.from_numpy(atm).to(dev)
    d_ndep = torch.full((ncol,), nd, dtype=torch.int32, device=dev)
    d_st = {p: torch.empty((ncol, 4, nout), dtype=torch.float64, device=dev) for p in (0, 1)}
    d_rf = {p: torch.empty((RF_NPARAM, nd, 4, nout, ncol),
                           dtype=torch.float32 if p else torch.float64, device=dev) for p in (0, 1)}
    torch.cuda.synchronize()

    def adjoint(p):
        eng.rf_lte_device(d_atm.data_ptr(), d_ndep.data_ptr(), ncol, nd, d_st[p].data_ptr(),
                          d_rf[p].data_ptr(), precision=p)

    t_adj, stages = {}, {}
    for p in (1, 0):
        adjoint(p)
        t0 = time.perf_counter()
        for _ in range(reps):
            adjoint(p)
        t_adj[p] = (time.perf_counter() - t0) / reps
        stages[p] = {k: round(float(v), 3) for k, v in eng.rf_timings().items()}

    blk = 16
    d_pert = torch.empty((blk * ncol, 9, nd), dtype=torch.float64, device=dev)
    d_pndep = torch.full((blk * ncol,), nd, dtype=torch.int32, device=dev)
    d_po = [torch.empty((blk * ncol, 4, nout), dtype=torch.float64, device=dev) for _ in (0, 1)]
    fd = torch.empty((RF_NPARAM, nd, 4, nout, ncol), dtype=torch.float64, device=dev)
    nsynth = 0

    def finite_differences():
        nonlocal nsynth
        nsynth = 0
        for ip, name in enumerate(RF_PARAMS):
            h, row = RF_STEP[name], RF_ROWS[ip]
            for k0 in range(0, nd, blk):
                ks = range(k0, min(nd, k0 + blk))
                nb = len(ks) * ncol
                for sg in (0, 1):
                    v = d_pert[:nb].view(len(ks), ncol, 9, nd)
                    v.copy_(d_atm.expand(len(ks), ncol, 9, nd))
                    for j, k in enumerate(ks):
                        v[j, :, row, k] += h if sg == 0 else -h
                    torch.cuda.current_stream().synchronize()
                    eng.synth_lte_device(d_pert.data_ptr(), d_pndep.data_ptr(), nb, nd,
                                         d_po[sg].data_ptr(), precision=0)
                    nsynth += len(ks)
                q = (d_po[0][:nb] - d_po[1][:nb]) / (2.0 * h)
                fd[ip, k0:k0 + len(ks)] = q.view(len(ks), ncol, 4, nout).permute(0, 2, 3, 1)
        torch.cuda.synchronize()

    t0 = time.perf_counter()
    finite_differences()
    t_fd = time.perf_counter() - t0

    # the conditioning reference: the FP64 adjoint of the same columns with
    # every input scaled by (1 + 1e-7 N(0, 1)), float's relative precision.
    # Its departures from the FP64 adjoint (points next to the kinks of the
    # limiter, clamps and Humlicek boundaries, where a rounding-size change
    # moves the derivative) are what the FP32 form's departures compare to.
    rng = np.random.default_rng(11)
    d_atm.mul_(torch.from_numpy(1.0 + 1e-7 * rng.standard_normal(atm.shape)).to(dev))
    d_rfp = torch.empty_like(d_rf[0])
    torch.cuda.synchronize()
    eng.rf_lte_device(d_atm.data_ptr(), d_ndep.data_ptr(), ncol, nd, d_po[0].data_ptr(),
                      d_rfp.data_ptr(), precision=0)
    cond = _err_stats(d_rfp, d_rf[0], 1e-4)
    del d_rfp

    I64 = d_st[0][:, 0]
    i_err = float(((d_st[1][:, 0] - I64).abs() / I64).max().item())
    stokes_err = float(((d_st[1] - d_st[0]).abs() / I64[:, :1, None]).max().item())
    res = {
        "workload": RF_WORKLOAD, "ncol": ncol, "ndep": nd, "nlambda": nout,
        "params": list(RF_PARAMS), "mu": 1.0,
        "adjoint_fp32": {"ms": round(t_adj[1] * 1e3, 2),
                         "columns_per_s": round(ncol / t_adj[1], 1), "stage_ms": stages[1]},
        "adjoint_fp64": {"ms": round(t_adj[0] * 1e3, 2),
                         "columns_per_s": round(ncol / t_adj[0], 1), "stage_ms": stages[0]},
        "fd_fp64": {"ms": round(t_fd * 1e3, 2), "columns_per_s": round(ncol / t_fd, 1),
                    "syntheses_per_column": nsynth, "steps": RF_STEP},
        "speedup_fp32_adjoint_vs_fd_fp64": round(t_fd / t_adj[1], 2),
        "speedup_fp64_adjoint_vs_fd_fp64": round(t_fd / t_adj[0], 2),
        "fp32_I_rel_err_max": float(f"{i_err:.3g}"),
        "fp32_stokes_err_over_Ic_max": float(f"{stokes_err:.3g}"),
        "rf_err_fp32_vs_fp64_adjoint": _err_stats(d_rf[1], d_rf[0], 1e-4),
        "rf_err_fp64_inputs_perturbed_1e-7_vs_fp64_adjoint": cond,
        "rf_err_fp64_adjoint_vs_fd": _err_stats(d_rf[0], fd, 1e-4),
        "rf_err_fp32_adjoint_vs_fd": _err_stats(d_rf[1], fd, 1e-4),
    }
    del d_rf, fd, d_pert, d_po
    torch.cuda.empty_cache()
    return res


def _nlte_leg(torch, dev, ncol, do_cpu, nworkers):
    """NLTE populations of the 4096-column batch: one rhg_nlte_solve of the
    whole batch with the atmospheres resident in HBM (set-up, scattering
    passes, iterations until every column has converged, J with the
    converged populations and the emergent ray at mu = 1, everything an
    rhf1d() call does), columns/s, the
    iteration histogram, the time split and the roofline of the formal
    solution kernel k_nlte_rt; the populations of 16 columns against the
    oracle; the reference's rhf1d() on a bounded sample of the same columns
    on this host's cores."""
    from pyrh_amd import model as M, nlte as NL, roofline as R, workloads as W
    from pyrh_amd.engine import Engine
    sys.path.insert(0, str(ROOT / "oracle"))
    wave = np.array(NLTE_WAVE)
    nm = NL.build_nlte(M.open_run(str(NLTE_RUN)), wave)
    atm = W.stokes_map(ncol, seed=W.SEED_BASE + 5)
    nd = atm.shape[2]
    eng = Engine(nm.lte)
    d_atm = torch.from_numpy(atm).to(dev)
    d_ndep = torch.full((ncol,), nd, dtype=torch.int32, device=dev)
    d_pops = torch.empty((ncol, nm.nlevel, nd), dtype=torch.float64, device=dev)
    d_niter = torch.empty(ncol, dtype=torch.int32, device=dev)
    d_dp = torch.empty(ncol, dtype=torch.float64, device=dev)
    d_st = torch.empty((ncol, 4, nm.lam.size - 1), dtype=torch.float64, device=dev)

    def solve(n):
        eng.nlte_solve_device(nm, d_atm.data_ptr(), d_ndep.data_ptr(), n, nd, d_pops.data_ptr(),
                              d_niter.data_ptr(), d_dp.data_ptr(), d_stokes=d_st.data_ptr())

    solve(min(ncol, 64))            # warm-up (module load, allocations)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solve(ncol)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    tm = eng.nlte_timings()
    niter = d_niter.cpu().numpy()
    pops = d_pops[:16].cpu().numpy()
    import pyoracle
    cp, cn, _ = pyoracle.nlte_solve(nm, atm[:16], nthreads=nworkers)
    err = float(np.max(np.abs(pops - cp) / cp))
    # the dominant kernel, k_nlte_bezier (wavelengths with a line): its
    # operations per column and pass from the census of the active sets,
    # times the columns of the passes without and with the operator
    cen = R.nlte_bezier_flops(nm, nd)
    flop_bez = cen["no_op"] * tm["cols_no_op"] + cen["op"] * tm["cols_op"]
    nset = np.diff(nm.as_ptr)
    nbez = int(np.sum((nset > 0) & (nm.as_bezier != 0)))
    pts = (tm["cols_no_op"] + tm["cols_op"]) * nbez * 2 * nm.muz.size * nd
    rt_ms = tm["scatter_bezier_ms"] + tm["iter_bezier_ms"]
    achieved = flop_bez / (rt_ms * 1e-3) / 1e12
    hist = {int(k): int(v) for k, v in zip(*np.unique(niter, return_counts=True))}
    res = {
        "workload": NLTE_WORKLOAD if ncol == 4096 else f"nlte_caii_{ncol}",
        "value": round(ncol / dt, 2), "unit": "columns/s", "wall_s": round(dt, 3),
        "dtype": "f64", "ncol": ncol, "ndep": nd, "nlevel": int(nm.nlevel),
        "ntrans": int(nm.tr_type.size), "nlambda": int(nm.lam.size), "nrays": int(nm.muz.size),
        "iter_limit": float(nm.iter_limit), "iterations_histogram": hist,
        "time_ms": {k: round(float(v), 3) for k, v in tm.items() if k.endswith("_ms")},
        "formal_passes": int(tm["rt_passes"]),
        "roofline": {"bound": "valu", "kernel": "k_nlte_bezier", "achieved": round(achieved, 3),
                     "peak": R.PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / R.PEAK_FP64_TFLOPS, 4), "traffic": None,
                     "flop_per_point": round(flop_bez / pts, 2),
                     "flop_per_column_pass": {"no_op": cen["no_op"], "op": cen["op"]},
                     "avg_launch_ms": round(rt_ms / max(tm["rt_passes"], 1), 3)},
        "oracle_check": {"columns": 16, "max_rel_err_pops": float(f"{err:.3g}"),
                         "iterations_equal": bool(np.array_equal(niter[:16], cn))},
        "cpu_baseline": None,
    }
    if do_cpu:
        import refrun
        ns = min(ncol, 2 * nworkers)
        if refrun.ref_available("o1"):
            from concurrent.futures import ThreadPoolExecutor
            chunks = np.array_split(atm[:ns], nworkers)
            t0 = time.perf_counter()
            with ThreadPoolExecutor(nworkers) as ex:
                list(ex.map(lambda x: refrun.run_synth(NLTE_RUN, x, wave, variant="o1"), chunks))
            dtc = time.perf_counter() - t0
            res["cpu_baseline"] = {
                "value": round(ns / dtc, 3), "unit": "columns/s", "cores": nworkers, "cpu": _host_cpu(),
                "kind": "reference",
                "sample": f"{ns} columns of the batch, rhf1d() per column to convergence plus its "
                          f"emergent spectrum, {nworkers} processes, {dtc:.1f} s wall"}
        else:
            t0 = time.perf_counter()
            pyoracle.nlte_solve(nm, atm[:ns], nthreads=nworkers)
            dtc = time.perf_counter() - t0
            res["cpu_baseline"] = {
                "value": round(ns / dtc, 3), "unit": "columns/s", "cores": nworkers, "cpu": _host_cpu(), "kind": "port",
                "sample": f"{ns} columns, C restatement, OpenMP {nworkers} threads, {dtc:.1f} s wall"}
    del d_atm, d_pops
    eng.close()
    torch.cuda.empty_cache()
    return res


HSE_WORKLOAD = "hse_map_16384"


def _hse_leg(torch, dev, ncol, do_cpu, nworkers, sample=128):
    """Hydrostatic equilibrium (pyrh's hse()) of the map's temperature
    columns: one rhg_hse call on ncol columns of 82 depths with the inputs
    resident in HBM, columns/s of the whole call and of the kernel, the
    pressure-iteration census and the roofline of k_hse (algorithmic FP64
    operations of the reference's arithmetic, pyrh_amd.roofline.hse_flops,
    with the Newton and chemical-equilibrium counts per pressure iteration
    from the oracle's census of `sample` columns); the GPU result on those
    columns against the oracle; the reference's hse() on a bounded sample on
    this host's cores."""
    from pyrh_amd import model as M, hse as HS, roofline as R, workloads as W
    from pyrh_amd.engine import Engine
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    hm = HS.build_hse(M.open_run(str(RUN_DIR)))
    T = np.ascontiguousarray(W.stokes_map(ncol, seed=W.SEED_BASE + 44)[:, 1])
    nd = T.shape[1]
    sc = np.ascontiguousarray(np.broadcast_to(W.LOGTAU, T.shape))
    eng = Engine(hm.model)
    d = {k: torch.from_numpy(v).to(dev) for k, v in
         (("sc", sc), ("T", T), ("pg", np.full(ncol, 0.1)), ("nd", np.full(ncol, nd, np.int32)))}
    outs = [torch.empty((ncol, nd), dtype=torch.float64, device=dev) for _ in range(4)]
    d_it = torch.empty((ncol, nd), dtype=torch.int32, device=dev)

    def run(n):
        eng.hse_device(hm, d["sc"].data_ptr(), d["T"].data_ptr(), d["pg"].data_ptr(),
                       d["nd"].data_ptr(), n, nd, *(o.data_ptr() for o in outs), 0,
                       d_it.data_ptr())

    run(min(ncol, 256))
    torch.cuda.synchronize()
    walls, kms = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        run(ncol)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
        kms.append(eng.hse_timing()[0])
    lanes = eng.hse_timing()[1]
    it = d_it.cpu().numpy()
    got = [o[:sample].cpu().numpy() for o in outs]
    *ref, rit, (newton, chem) = pyoracle.hse(hm, sc[:sample], T[:sample], 0.1, nthreads=nworkers,
                                             census=True)
    err = max(float(np.max(np.abs(a - b) / np.abs(b))) for a, b in zip(got, ref))
    L_s = int(rit.sum())
    fl, per = R.hse_flops(hm, int(it.sum()), newton / L_s, chem / L_s)
    k_ms = min(kms)
    achieved = fl / (k_ms * 1e-3) / 1e12
    res = {
        "workload": HSE_WORKLOAD if ncol == 16384 else f"hse_map_{ncol}",
        "value": round(ncol / min(walls), 1), "unit": "columns/s",
        "kernel_columns_per_s": round(ncol / (k_ms * 1e-3), 1), "kernel_ms": round(k_ms, 3),
        "wall_ms": round(min(walls) * 1e3, 3), "dtype": "f64", "ncol": ncol, "ndep": nd,
        "lanes_per_column": lanes, "pressure_iterations": int(it.sum()),
        "top_layer_iterations_mean": round(float(it[:, 0].mean()), 2),
        "census_per_pressure_iteration": {"newton_steps": round(newton / L_s, 4),
                                          "chem_iterations": round(chem / L_s, 4)},
        "roofline": {"bound": "valu", "kernel": "k_hse", "achieved": round(achieved, 3),
                     "peak": R.PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / R.PEAK_FP64_TFLOPS, 4), "traffic": None,
                     "flops_per_launch": fl, "flop_model": per},
        "oracle_check": {"columns": sample, "max_rel_err": float(f"{err:.3g}"),
                         "iterations_equal": bool(np.array_equal(it[:sample], rit))},
        "cpu_baseline": None,
    }
    if do_cpu:
        import refrun
        ns = min(ncol, 64 * nworkers)
        if refrun.ref_available("o1"):
            from concurrent.futures import ThreadPoolExecutor
            chunks = np.array_split(np.arange(ns), nworkers)
            t0 = time.perf_counter()
            with ThreadPoolExecutor(nworkers) as ex:
                list(ex.map(lambda ix: refrun.run_hse(RUN_DIR, W.LOGTAU, T[ix], variant="o1"),
                            chunks))
            dtc = time.perf_counter() - t0
            res["cpu_baseline"] = {
                "value": round(ns / dtc, 3), "unit": "columns/s", "cores": nworkers, "cpu": _host_cpu(),
                "kind": "reference",
                "sample": f"{ns} columns of the batch, hse() per column, {nworkers} processes, "
                          f"{dtc:.1f} s wall"}
        else:
            t0 = time.perf_counter()
            pyoracle.hse(hm, sc[:ns], T[:ns], 0.1, nthreads=nworkers)
            dtc = time.perf_counter() - t0
            res["cpu_baseline"] = {
                "value": round(ns / dtc, 3), "unit": "columns/s", "cores": nworkers, "cpu": _host_cpu(), "kind": "port",
                "sample": f"{ns} columns, C restatement, OpenMP {nworkers} threads, {dtc:.1f} s wall"}
    del outs, d
    eng.close()
    torch.cuda.empty_cache()
    return res


MIXED_RUN = ROOT / "tests" / "golden" / "run_mixed"
MIXED_WORKLOAD = "mixed_zeeman400_8192"


def _mixed_leg(torch, dev, ncol, do_cpu, nworkers, rf_cols=256, oracle_cols=128):
    """The mixed-size batch: 8192 columns with 64-256 depth points (seeded),
    400 synthetic Fe I Zeeman lines on 2000 wavelengths, FP32 Stokes at
    mu = 1, 0.7, 0.4 with the batch resident in HBM (one rhg_synth_lte per
    mu; the library runs the batch sorted by depth count); FP32 adjoint
    response functions to T, v_LOS and B of all columns at mu = 1, in chunks
    of rf_cols columns (the tape of 2000 wavelengths x 256 depths takes
    ~0.2 GB per column); the FP64 oracle on oracle_cols columns at mu = 0.7; the
    reference's rhf1d() per (column, mu) on a bounded sample.  The unit is
    (column, mu) syntheses per second."""
    from pyrh_amd import model as M, roofline as R, workloads as W
    from pyrh_amd.engine import Engine
    sys.path.insert(0, str(ROOT / "oracle"))
    wave = W.mixed_wavelengths()
    mdl = M.build_model(M.open_run(str(MIXED_RUN)), wave)
    atm, nd = W.mixed_batch(ncol, seed=W.SEED_BASE + 8)
    nmax = atm.shape[2]
    eng = Engine(mdl)
    nout = mdl.scalars["nspect"] - 1
    d_atm = torch.from_numpy(atm).to(dev)
    d_nd = torch.from_numpy(nd).to(dev)
    d_out = torch.empty((len(W.MIXED_MU), ncol, 4, nout), dtype=torch.float64, device=dev)

    def synth(prec):
        for i, mu in enumerate(W.MIXED_MU):
            eng.synth_lte_device(d_atm.data_ptr(), d_nd.data_ptr(), ncol, nmax, d_out[i].data_ptr(),
                                 mu=mu, precision=prec)

    # launch configuration (pyrh_amd.autotune) on the first 2048 columns,
    # FP32 form, mu = 1; it applies to the whole batch and every mu
    from pyrh_amd import autotune
    nt = min(ncol, 2048)
    tuning = autotune.tune(eng, d_atm[:nt], d_nd[:nt], d_out[0, :nt], precision=1, reps=1)

    synth(1)                       # warm-up (module load, allocations)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    synth(1)
    torch.cuda.synchronize()
    dt32 = time.perf_counter() - t0
    formal_ms = eng.timings()["formal"]
    fl = R.flops(eng.count_work(), mdl.scalars["nline"])
    st32 = d_out[1, :oracle_cols].cpu().numpy()
    t0 = time.perf_counter()
    synth(0)
    torch.cuda.synchronize()
    dt64 = time.perf_counter() - t0
    st64 = d_out[1, :oracle_cols].cpu().numpy()
    import pyoracle
    ref = pyoracle.synth_lte(mdl, atm[:oracle_cols], mu=W.MIXED_MU[1], ndep=nd[:oracle_cols],
                             nthreads=nworkers)
    Ic = ref[:, 0, :1]
    achieved = fl["formal_kernel"] / (formal_ms * 1e-3) / 1e12
    res = {
        "workload": MIXED_WORKLOAD if ncol == 8192 else f"mixed_zeeman400_{ncol}",
        "unit": "(column, mu) syntheses/s", "ncol": ncol, "mu": list(W.MIXED_MU),
        "ndep_counts": {int(k): int(v) for k, v in zip(*np.unique(nd, return_counts=True))},
        "nlambda": nout, "nline": int(mdl.scalars["nline"]),
        "autotune": tuning,
        "fp32": {"value": round(len(W.MIXED_MU) * ncol / dt32, 1), "wall_s": round(dt32, 3)},
        "fp64": {"value": round(len(W.MIXED_MU) * ncol / dt64, 1), "wall_s": round(dt64, 3)},
        "roofline_fp32_mu04": {"kernel": "k_formal_stokes<float>", "achieved": round(achieved, 3),
                               "peak": R.PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                               "frac": round(achieved / R.PEAK_FP32_TFLOPS, 4),
                               "flops_per_launch": fl["formal_kernel"],
                               "launch_ms": round(formal_ms, 3)},
        "oracle_check_mu07": {
            "columns": oracle_cols,
            "fp64_I_rel_err_max": float(f"{np.max(np.abs(st64[:, 0] - ref[:, 0]) / ref[:, 0]):.3g}"),
            "fp64_QUV_err_over_Ic_max": float(f"{np.max(np.abs(st64[:, 1:] - ref[:, 1:]) / Ic[:, None]):.3g}"),
            "fp32_I_err_over_Ic_max": float(f"{np.max(np.abs(st32[:, 0] - ref[:, 0]) / Ic):.3g}"),
            "fp32_QUV_err_over_Ic_max": float(f"{np.max(np.abs(st32[:, 1:] - ref[:, 1:]) / Ic[:, None]):.3g}")},
    }
    del d_out
    torch.cuda.empty_cache()
    # FP32 adjoint response functions to T, v_LOS and B of every column at
    # mu = 1, in chunks of rf_cols columns ordered by depth count (each
    # chunk's output, 3 x ndep x 4 x 2000 x rf_cols floats, is consumed and
    # its buffer reused); the first chunk also in FP64 for the error
    params = ("T", "vlos", "B")
    order = np.argsort(-nd, kind="stable")
    nrf = min(rf_cols, ncol)
    d_st = torch.empty((nrf, 4, nout), dtype=torch.float64, device=dev)
    d_rf = torch.empty((len(params), nmax, 4, nout, nrf), dtype=torch.float32, device=dev)

    def written(buf, n, m):
        # the library's layout for n columns of ndep_max m
        return buf.view(-1)[:len(params) * m * 4 * nout * n].view(len(params), m, 4, nout, n)

    def chunk(c0, prec=1, out=None):
        idx = order[c0:c0 + nrf]
        n, m = len(idx), int(nd[idx].max())
        a = torch.from_numpy(np.ascontiguousarray(atm[idx][:, :, :m])).to(dev)
        dn = torch.from_numpy(nd[idx]).to(dev)
        torch.cuda.synchronize()
        eng.rf_lte_device(a.data_ptr(), dn.data_ptr(), n, m, d_st.data_ptr(),
                          (d_rf if out is None else out).data_ptr(), mu=1.0, precision=prec,
                          params=params)
        return n, m

    chunk(0)                       # warm-up
    torch.cuda.synchronize()
    stages = {k: 0.0 for k in eng.rf_timings()}
    t0 = time.perf_counter()
    for c0 in range(0, ncol, nrf):
        n, m = chunk(c0)
        for k, v in eng.rf_timings().items():
            stages[k] += float(v)
    torch.cuda.synchronize()
    dtr = time.perf_counter() - t0
    # the last chunk (shallowest columns) is still in the buffer
    finite = bool(torch.isfinite(written(d_rf, n, m)).all().item())
    n0, m0 = chunk(0)
    torch.cuda.synchronize()
    rf32 = written(d_rf, n0, m0).clone()
    d_rf64 = torch.empty((len(params), m0, 4, nout, n0), dtype=torch.float64, device=dev)
    chunk(0, prec=0, out=d_rf64)
    torch.cuda.synchronize()
    err = _err_stats(rf32, d_rf64, 1e-4, params)
    finite &= bool(torch.isfinite(rf32).all().item())
    res["adjoint_rf_fp32"] = {"workload": f"{res['workload']}_rf_T_vlos_B", "columns": ncol,
                              "mu": 1.0, "params": list(params), "chunk_columns": nrf,
                              "ms": round(dtr * 1e3, 2), "columns_per_s": round(ncol / dtr, 1),
                              "stage_ms_total": {k: round(v, 2) for k, v in stages.items()},
                              "finite": finite,
                              "err_vs_fp64_adjoint_first_chunk": err}
    del d_rf, d_st, d_rf64, rf32
    eng.close()
    torch.cuda.empty_cache()
    res["cpu_baseline"] = None
    if do_cpu:
        import refrun
        if refrun.ref_available("o1"):
            from concurrent.futures import ThreadPoolExecutor
            sample = [(c, mu) for c in range(nworkers) for mu in W.MIXED_MU[:1]]
            t0 = time.perf_counter()
            with ThreadPoolExecutor(nworkers) as ex:
                list(ex.map(lambda cm: refrun.run_synth(MIXED_RUN, atm[cm[0]:cm[0] + 1, :, :nd[cm[0]]],
                                                        wave, mu=cm[1], variant="o1"), sample))
            dtc = time.perf_counter() - t0
            res["cpu_baseline"] = {
                "value": round(len(sample) / dtc, 3), "unit": "(column, mu) syntheses/s",
                "cores": nworkers, "cpu": _host_cpu(), "kind": "reference",
                "sample": f"{len(sample)} columns of the batch (depths {sorted(set(nd[:nworkers].tolist()))}) "
                          f"at mu = 1, rhf1d() per column with its input parsing, {nworkers} processes, "
                          f"{dtc:.1f} s wall"}
    return res


def _grids_leg(torch, dev, setup, wave, atm, reps=10, ngrid=4):
    """Uneven wavelength grids in one launch: the LTE map's columns, column i
    asking for grid i mod 4 of the 112 wavelengths cut into four windows of
    28, as one rhg_synth_lte_grids call on the merged context, against four
    rhg_synth_lte calls on contexts built on one window each (the columns
    of each window already gathered on the device).  FP32 form; both sides
    bit-identical on every column."""
    from pyrh_amd import api, model as M
    from pyrh_amd.engine import Engine
    ncol, nd = atm.shape[0], atm.shape[2]
    n = wave.size
    waves = [wave[g * n // ngrid:(g + 1) * n // ngrid] for g in range(ngrid)]
    e = Engine(M.build_model(setup, np.unique(np.concatenate(waves))))
    pos = api.grid_positions(e.model, waves)
    e.set_grids(pos)
    singles = [Engine(M.build_model(setup, w)) for w in waves]
    gid = (np.arange(ncol) % ngrid).astype(np.int32)
    parts = [np.flatnonzero(gid == g) for g in range(ngrid)]
    d_atm = torch.from_numpy(atm).to(dev)
    d_nd = torch.full((ncol,), nd, dtype=torch.int32, device=dev)
    d_g = torch.from_numpy(gid).to(dev)
    d_out = torch.empty((ncol, 4, e.nout), dtype=torch.float64, device=dev)
    d_atm_g = [d_atm[torch.from_numpy(p).to(dev)].contiguous() for p in parts]
    d_out_g = [torch.empty((p.size, 4, s.nout), dtype=torch.float64, device=dev)
               for p, s in zip(parts, singles)]

    def one():
        e.synth_lte_grids_device(d_atm.data_ptr(), d_nd.data_ptr(), d_g.data_ptr(), ncol, nd,
                                 d_out.data_ptr(), precision=1)

    def split():
        for g, s in enumerate(singles):
            s.synth_lte_device(d_atm_g[g].data_ptr(), d_nd.data_ptr(), parts[g].size, nd,
                               d_out_g[g].data_ptr(), precision=1)

    res = {}
    for name, f in (("one_launch", one), ("per_grid_calls", split)):
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / reps
    out = d_out.cpu().numpy()
    same = all(np.array_equal(out[p][:, :, pos[g]], d_out_g[g].cpu().numpy())
               for g, p in enumerate(parts))
    for s in singles:
        s.close()
    e.close()
    return {"workload": f"lte_stokes_map_{ncol}_{ngrid}_grids",
            "grids": [int(p.size) for p in pos],
            "one_launch_ms": round(res["one_launch"] * 1e3, 3),
            "per_grid_calls_ms": round(res["per_grid_calls"] * 1e3, 3),
            "value": round(ncol / res["one_launch"], 1), "unit": "columns/s",
            "bit_identical": bool(same)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ncol", type=int, default=16384)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rf", action="store_true", help="skip the response-function leg")
    ap.add_argument("--no-nlte", action="store_true", help="skip the NLTE leg")
    ap.add_argument("--nlte-ncol", type=int, default=4096)
    ap.add_argument("--nlte-only", action="store_true",
                    help="run only the NLTE leg and print its object (for profiling)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-size batch leg")
    ap.add_argument("--mixed-ncol", type=int, default=8192)
    ap.add_argument("--mixed-only", action="store_true",
                    help="run only the mixed-size batch leg and print its object")
    ap.add_argument("--no-hse", action="store_true", help="skip the hydrostatic-equilibrium leg")
    ap.add_argument("--no-grids", action="store_true", help="skip the wavelength-grids leg")
    ap.add_argument("--hse-ncol", type=int, default=16384)
    ap.add_argument("--hse-only", action="store_true",
                    help="run only the hydrostatic-equilibrium leg and print its object")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="each kernel of the bench once at its workload shape and one 1 GiB "
                         "device copy, for rocprofv3 --pmc")
    ap.add_argument("--profile-step", action="store_true",
                    help="one FP32-form synthesis and one 1 GiB device copy, for rocprofv3 --pmc")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); columns are "
              f"independent and a multi-GPU run would shard them, which is not built.",
              file=sys.stderr)
        sys.exit(2)

    nw = max(1, min(16, len(os.sched_getaffinity(0))))
    if a.profile_kernels:
        _profile_kernels()
        return
    if a.nlte_only or a.mixed_only or a.hse_only:
        import torch
        leg, n = ((_nlte_leg, a.nlte_ncol) if a.nlte_only else
                  (_mixed_leg, a.mixed_ncol) if a.mixed_only else (_hse_leg, a.hse_ncol))
        res = leg(torch, torch.device("cuda:0"), n, not a.no_cpu, nw)
        print(json.dumps(res), flush=True)
        return

    from pyrh_amd import model as M, roofline as R, workloads as W
    setup = M.open_run(str(RUN_DIR))
    wave = W.wavelengths_630()
    mdl = M.build_model(setup, wave)
    atm = W.stokes_map(a.ncol, seed=W.SEED_BASE + 1)
    ndep = np.full(a.ncol, atm.shape[2], np.int32)

    import torch
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a HIP device")
    from pyrh_amd.engine import Engine
    dev = torch.device("cuda:0")
    d_atm = torch.from_numpy(atm).to(dev)
    d_ndep = torch.from_numpy(ndep).to(dev)
    nout = mdl.scalars["nspect"] - 1
    d_out = torch.empty((a.ncol, 4, nout), dtype=torch.float64, device=dev)
    eng = Engine(mdl, device=0)

    def step(precision=0):
        eng.synth_lte_device(d_atm.data_ptr(), d_ndep.data_ptr(), a.ncol, atm.shape[2],
                             d_out.data_ptr(), precision=precision)

    if a.profile_step:
        step(1)
        x = torch.empty(1 << 30, dtype=torch.uint8, device=dev)
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        print(f"profile step done: {a.ncol} columns; 1 GiB copy", flush=True)
        return

    # launch configuration for this batch (pyrh_amd.autotune), FP32 form
    from pyrh_amd import autotune
    tuning = autotune.tune(eng, d_atm, d_ndep, d_out, precision=1, reps=3)

    def timed(precision):
        for _ in range(a.warmup):
            step(precision)
        torch.cuda.synchronize()
        stage_ms = []
        t0 = time.perf_counter()
        for _ in range(a.steps):
            step(precision)
            stage_ms.append(eng.timings())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        stages = {k: round(float(np.mean([s[k] for s in stage_ms])), 3) for k in stage_ms[0]}
        return dt, stages

    dt64, stages64 = timed(0)
    fl = R.flops(eng.count_work(), mdl.scalars["nline"])
    dt, stages = timed(1)
    ms_step = dt / a.steps * 1e3
    value = a.ncol * a.steps / dt
    formal_ms = stages["formal"]
    achieved = fl["formal_kernel"] / (formal_ms * 1e-3) / 1e12
    ach64 = fl["formal_kernel"] / (stages64["formal"] * 1e-3) / 1e12

    rf = None if a.no_rf else _rf_leg(eng, W, torch, dev, reps=3)
    nlte = None if a.no_nlte else _nlte_leg(torch, dev, a.nlte_ncol, not a.no_cpu, nw)
    mixed = None if a.no_mixed else _mixed_leg(torch, dev, a.mixed_ncol, not a.no_cpu, nw)
    hse = None if a.no_hse else _hse_leg(torch, dev, a.hse_ncol, not a.no_cpu, nw)
    grids = None if a.no_grids else _grids_leg(torch, dev, setup, wave, atm)

    # host-buffer rate (PCIe both ways), reported beside the resident rate
    t1 = time.perf_counter()
    eng.synth_lte(atm, ndep=ndep, precision=1)
    host_rate = a.ncol / (time.perf_counter() - t1)

    cpu = None
    if not a.no_cpu:
        cpu = _cpu_baseline(setup, wave, atm, nw)

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "columns/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded LTE Stokes map; line and atom data from the run directory)",
        "config": {"workload": WORKLOAD if a.ncol == 16384 else f"lte_stokes_map_{a.ncol}",
                   "ncol": a.ncol, "ndep": int(atm.shape[2]), "nlambda": nout, "mu": 1.0,
                   "lines": "Fe I 630.15 + 630.25 nm, Zeeman",
                   "form": "FP32 form: float continuum and Voigt/Faraday-Voigt functions; FP64 "
                           "Zeeman sums, points, DELO-Bezier recursion, equilibria and heights",
                   "response_functions": RF_WORKLOAD + " (object response_functions)"},
        "roofline": {"bound": "valu", "kernel": "k_formal_stokes<float>",
                     "achieved": round(achieved, 3), "peak": R.PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / R.PEAK_FP32_TFLOPS, 4),
                     "traffic": _traffic_from_profile(a.ncol),
                     "flops_per_launch": fl["formal_kernel"], "launch_ms": round(formal_ms, 3)},
        "cpu_baseline": cpu,
        "stage_ms": stages,
        "fp64_plain": {"value": round(a.ncol * a.steps / dt64, 1), "unit": "columns/s",
                       "ms_per_step": round(dt64 / a.steps * 1e3, 3), "stage_ms": stages64,
                       "roofline": {"kernel": "k_formal_stokes<double>",
                                    "achieved": round(ach64, 3), "peak": R.PEAK_FP64_TFLOPS,
                                    "unit": "TFLOP/s",
                                    "frac": round(ach64 / R.PEAK_FP64_TFLOPS, 4)}},
        "response_functions": rf,
        "nlte": nlte,
        "mixed": mixed,
        "hse": hse,
        "grids": grids,
        "autotune": tuning,
        "pipeline_flops": fl["total"],
        "host_buffers_columns_per_s": round(host_rate, 1),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
