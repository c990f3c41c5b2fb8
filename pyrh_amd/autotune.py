"""Autotuner of a context's launch configuration.

The library's kernels have a few launch-configuration choices that change
their speed but not their results (include/pyrh_amd.h, rhg_set_tuning):

* line_table:  stage each wavelength's line list (lines, Doppler offsets,
  Zeeman components) in LDS once per wavefront for the polarized formal
  solution, or read it from the model in global memory at every depth;
* column_sort: run a batch of uneven depth counts sorted by depth count
  (a wavefront of 64 columns waits for its deepest column), or as given;
* wave_order:  launch the polarized wavelengths in wavelength order, or
  those with the most Zeeman components first (the longest workgroups
  start first and the tail of the launch is shorter).

`tune(engine, atmos, ndep, out, ...)` times every combination on the
caller's own batch (CUDA tensors, as for Engine.synth_lte_device), keeps
the fastest, checks that every combination gave the same Stokes vectors bit
for bit, and returns a report of the candidates' times.  The choice depends
on the model (how many lines and components a wavelength's list holds), the
batch (its depth counts) and the precision form, so the tuner runs once per
such combination; `cached_tune` keeps the result per engine and batch
shape.
"""
from __future__ import annotations

import itertools
import time

import numpy as np

from .engine import TUNE_KEYS

DEFAULTS = {"line_table": 1, "column_sort": 1, "wave_order": 0}


def tune(engine, atmos, ndep, out, mu=1.0, atm_scale=0, precision=1, reps=3, knobs=TUNE_KEYS):
    """Time every combination of `knobs` (the others at their defaults) on
    the batch atmos [ncol, 9, ndep_max] (CUDA float64), ndep [ncol] (CUDA
    int32) with Engine.synth_lte_device into out [ncol, 4, nout] (CUDA
    float64) and leave the fastest set on the engine.  Each candidate's
    Stokes vectors are compared with the first's, bit for bit (RuntimeError
    otherwise).  Returns {"chosen": {knob: value}, "candidates": [{knob:
    value, "ms": median wall time of the call, "formal_ms": the formal
    solution's kernel time}], "ncol", "ndep_max", "precision", "reps"}."""
    import torch
    ncol, _, ndep_max = atmos.shape

    def call():
        engine.synth_lte_device(atmos.data_ptr(), ndep.data_ptr(), ncol, ndep_max,
                                out.data_ptr(), mu=mu, atm_scale=atm_scale, precision=precision)

    torch.cuda.synchronize()
    cands, first = [], None
    for values in itertools.product((0, 1), repeat=len(knobs)):
        cfg = dict(DEFAULTS)
        cfg.update(zip(knobs, values))
        for k in TUNE_KEYS:
            engine.set_tuning(k, cfg[k])
        call()                                  # warm-up under this setting
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            call()
            ts.append((time.perf_counter() - t0) * 1e3)
        formal = float(engine.timings()["formal"])
        if first is None:
            first = out.clone()
        elif not torch.equal(first, out):
            raise RuntimeError(f"launch configuration {cfg} changed the Stokes vectors")
        cands.append({**engine.tuning(), "ms": round(float(np.median(ts)), 3),
                      "formal_ms": round(formal, 3)})
    best = min(cands, key=lambda c: c["ms"])
    for k in TUNE_KEYS:
        engine.set_tuning(k, best[k])
    return {"chosen": {k: best[k] for k in TUNE_KEYS}, "candidates": cands, "ncol": int(ncol),
            "ndep_max": int(ndep_max), "precision": int(precision), "reps": int(reps)}


_CACHE = {}


def cached_tune(engine, atmos, ndep, out, **kw):
    """tune() once per (engine, batch shape, precision, mu); later calls
    re-apply the stored choice."""
    key = (id(engine), tuple(atmos.shape), int(kw.get("precision", 1)), float(kw.get("mu", 1.0)))
    rep = _CACHE.get(key)
    if rep is None:
        rep = tune(engine, atmos, ndep, out, **kw)
        _CACHE[key] = rep
    else:
        for k, v in rep["chosen"].items():
            engine.set_tuning(k, v)
    return rep
