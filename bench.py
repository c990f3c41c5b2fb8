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
bounded sample of the same columns.

    python bench.py [--steps K] [--warmup W] [--ncol N] [--no-cpu]
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


def _traffic_from_profile(ncol):
    """HBM bytes per k_formal launch from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    f = PROFILES / "k_formal_lte_stokes_map_16384_hbm.md"
    if not f.exists() or ncol != 16384:
        return None
    m = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", f.read_text())
    return float(m.group(1)) if m else None


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
        return {"value": ns / dt, "unit": "columns/s", "cores": nworkers, "kind": "reference",
                "sample": f"{ns} columns of the map, rhf1d() per column (input files re-read "
                          f"per call as pyrh does), {nworkers} processes, {dt:.1f} s wall"}
    import pyoracle
    from pyrh_amd import model as M
    m = M.build_model(setup, wave)
    t0 = time.perf_counter()
    pyoracle.synth_lte(m, sample, nthreads=nworkers)
    dt = time.perf_counter() - t0
    return {"value": ns / dt, "unit": "columns/s", "cores": nworkers, "kind": "port",
            "sample": f"{ns} columns of the map, C restatement, OpenMP {nworkers} threads, "
                      f"{dt:.1f} s wall"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ncol", type=int, default=16384)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-step", action="store_true",
                    help="one FP32-form synthesis and one 1 GiB device copy, for rocprofv3 --pmc")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); columns are "
              f"independent and a multi-GPU run would shard them, which is not built.",
              file=sys.stderr)
        sys.exit(2)

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

    # host-buffer rate (PCIe both ways), reported beside the resident rate
    t1 = time.perf_counter()
    eng.synth_lte(atm, ndep=ndep, precision=1)
    host_rate = a.ncol / (time.perf_counter() - t1)

    cpu = None
    if not a.no_cpu:
        nw = max(1, min(16, len(os.sched_getaffinity(0))))
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
                   "form": "FP32 form: float line opacity and DELO-Bezier recursion; FP64 "
                           "equilibria, heights, path lengths and Bezier weights",
                   "response_functions": "not built"},
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
        "pipeline_flops": fl["total"],
        "host_buffers_columns_per_s": round(host_rate, 1),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
