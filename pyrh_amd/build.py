"""Build the gfx950 HIP library in-tree (pyrh_amd/libpyrh_amd.so)."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libpyrh_amd.so"
# translation units; the FP32 form's kernels are compiled with FMA
# contraction, everything else keeps the reference's operation order
SOURCES = ["rhg_api.hip", "rhg_kernels.hip", "rhg_formal.hip", "rhg_formal32.hip", "rhg_rf.hip",
           "rhg_rf32.hip", "rhg_nlte.hip", "rhg_hse.hip"]
CONTRACT = {"rhg_formal32.hip", "rhg_rf32.hip"}

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build pyrh_amd)")


def _jobs():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 8, len(SOURCES)))


def build_library(verbose: bool = False) -> Path:
    """Compile every HIP source (in parallel) and link LIB; always
    recompiles."""
    cc = hipcc()
    with tempfile.TemporaryDirectory(prefix="pyrh_amd_build_") as td:
        objs, procs = [], []
        pending = list(SOURCES)
        log = []

        def reap(block):
            for p in list(procs):
                if block or p[0].poll() is not None:
                    out, err = p[0].communicate()
                    procs.remove(p)
                    if p[0].returncode != 0:
                        raise RuntimeError(f"HIP build of {p[1]} failed:\n" + out + err)
                    log.append(err)

        while pending or procs:
            while pending and len(procs) < _jobs():
                src = pending.pop(0)
                obj = os.path.join(td, src.replace(".hip", ".o"))
                fpc = "-ffp-contract=fast" if src in CONTRACT else "-ffp-contract=off"
                cmd = [cc, *HIPCC_FLAGS, fpc, "-c", str(CSRC / src), "-o", obj]
                procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                               text=True), src))
                objs.append(obj)
            if procs:
                procs[0][0].wait()
                reap(False)
        tmp = LIB.with_suffix(".so.tmp")
        cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp)] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("HIP link failed:\n" + r.stdout + r.stderr)
    if verbose:
        sys.stderr.write("".join(log) + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
