"""Run the reference RH library built from source (oracle/_ref) on columns.

Test infrastructure only: fixture generation and the CPU baseline of the
benchmark.  The binary layouts are those documented at the top of
oracle/ref_harness.c.
"""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[1]
REF_DIR = REPO / "oracle" / "_ref"
DATA_DIR = REPO / "tests" / "golden" / "rhdata"
MAGIC = 0x4E494852


def ref_available(variant: str = "o2") -> bool:
    return (REF_DIR / f"rhref_run_{variant}").exists()


def _write_input(path, atmos, wave, mu, atm_scale):
    atmos = np.ascontiguousarray(atmos, dtype=np.float64)
    if atmos.ndim == 2:
        atmos = atmos[None]
    ncol, nrow, ndep = atmos.shape
    assert nrow == 9
    wave = np.ascontiguousarray(wave, dtype=np.float64)
    with open(path, "wb") as f:
        f.write(struct.pack("<5i", MAGIC, ncol, ndep, wave.size, atm_scale))
        f.write(struct.pack("<d", float(mu)))
        f.write(wave.tobytes())
        f.write(atmos.tobytes())
    return ncol, ndep


def read_capture(path):
    """Return a list of (tag, col, nspect, mu, to_obs, data) records."""
    out = []
    raw = Path(path).read_bytes()
    pos = 0
    while pos < len(raw):
        tag = raw[pos:pos + 16].split(b"\0", 1)[0].decode()
        col, nspect, mu, to_obs, n = struct.unpack_from("<5i", raw, pos + 16)
        pos += 36
        data = np.frombuffer(raw, dtype=np.float64, count=n, offset=pos).copy()
        pos += 8 * n
        out.append((tag, col, nspect, mu, to_obs, data))
    return out


def run_synth(rundir, atmos, wave, mu=1.0, atm_scale=0, variant="o2",
              capture=False, timeout=3600, loggf=(), lam=(), abun=(), rf=False, fudge=None):
    """Call rhf1d() once per column.  Returns (lam, I, Q, U, V[, capture]
    [, rf]).  loggf / lam / abun are sequences of (index, value) pairs passed
    as pyrh.compute1d's loggf_ids/values, lam_ids/values and
    atomic_number/abundance; rf requests get_atomic_rfs and returns the
    response functions [ncol][nlw][len(loggf) + len(lam)]."""
    exe = REF_DIR / f"rhref_run_{variant}"
    env = dict(os.environ, PYRH_PATH=str(DATA_DIR))
    with tempfile.TemporaryDirectory() as td:
        fin, fout, fcap = (os.path.join(td, n) for n in ("in", "out", "cap"))
        frf = os.path.join(td, "rf")
        npar = len(loggf) + len(lam)
        if loggf or lam or abun or rf or fudge is not None:
            fov = os.path.join(td, "overrides")
            with open(fov, "w") as f:
                for i, v in loggf:
                    f.write(f"loggf {int(i)} {float(v)!r}\n")
                for i, v in lam:
                    f.write(f"lam {int(i)} {float(v)!r}\n")
                for i, v in abun:
                    f.write(f"abun {int(i)} {float(v)!r}\n")
                if rf:
                    f.write(f"rf {frf}\n")
                if fudge is not None:
                    fw, fv = fudge
                    for k in range(len(fw)):
                        f.write(f"fudge {float(fw[k])!r} {float(fv[0][k])!r} "
                                f"{float(fv[1][k])!r} {float(fv[2][k])!r}\n")
            env["RHREF_OVERRIDES"] = fov
        ncol, ndep = _write_input(fin, atmos, wave, mu, atm_scale)
        cmd = [str(exe), str(rundir), fin, fout] + ([fcap] if capture else [])
        r = subprocess.run(cmd, env=env, capture_output=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"reference run failed ({r.returncode}): "
                               f"{r.stderr.decode()[-2000:]}")
        raw = Path(fout).read_bytes()
        pos = 0
        res = []
        for _ in range(ncol):
            (nlw,) = struct.unpack_from("<i", raw, pos)
            pos += 4
            arr = np.frombuffer(raw, np.float64, 5 * nlw, pos).reshape(5, nlw)
            pos += 40 * nlw
            res.append(arr.copy())
        res = np.stack(res)
        lam = res[0, 0]
        out = (lam, res[:, 1], res[:, 2], res[:, 3], res[:, 4])
        if capture:
            out = out + (read_capture(fcap),)
        if rf:
            out = out + (np.fromfile(frf, np.float64).reshape(ncol, -1, npar),)
        return out


def run_hse(rundir, scale, T, pg_top=0.1, variant="o2", timeout=600, atm_scale=0,
            abun=(), fudge=None):
    """pyrh.hse() once per column of T [ncol, ndep] on `scale` (log10 tau500
    for atm_scale 0, height [km] for 2; 1-D or per column); returns ne,
    nHtot, rho, pg [ncol, 4, ndep] exactly as the C entry point leaves them
    (SI units of pyrh_hse.c).  abun: (atomic number, abundance) pairs;
    fudge: (fudge_wave [n], fudge_value [3][n])."""
    exe = REF_DIR / f"rhref_run_{variant}"
    env = dict(os.environ, PYRH_PATH=str(DATA_DIR))
    T = np.atleast_2d(np.asarray(T, np.float64))
    ncol, ndep = T.shape
    atmos = np.zeros((ncol, 9, ndep))
    atmos[:, 0] = scale
    atmos[:, 1] = T
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in"), os.path.join(td, "out")
        if abun or fudge is not None:
            fov = os.path.join(td, "overrides")
            with open(fov, "w") as f:
                for i, v in abun:
                    f.write(f"abun {int(i)} {float(v)!r}\n")
                if fudge is not None:
                    fw, fv = fudge
                    for k in range(len(fw)):
                        f.write(f"fudge {float(fw[k])!r} {float(fv[0][k])!r} "
                                f"{float(fv[1][k])!r} {float(fv[2][k])!r}\n")
            env["RHREF_OVERRIDES"] = fov
        _write_input(fin, atmos, np.zeros(1), pg_top, atm_scale)
        r = subprocess.run([str(exe), "--hse", str(rundir), fin, fout],
                           env=env, capture_output=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"reference hse failed: {r.stderr.decode()[-2000:]}")
        o = np.fromfile(fout, np.float64).reshape(ncol, 4, ndep)
    return o


def _run_mode(mode, rundir, atmos, mu, atm_scale, nout, timeout=600, abun=()):
    exe = REF_DIR / "rhref_run_o2"
    env = dict(os.environ, PYRH_PATH=str(DATA_DIR))
    atmos = np.ascontiguousarray(atmos, np.float64)
    if atmos.ndim == 2:
        atmos = atmos[None]
    ncol, _, ndep = atmos.shape
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in"), os.path.join(td, "out")
        if abun:
            fov = os.path.join(td, "overrides")
            with open(fov, "w") as f:
                for i, v in abun:
                    f.write(f"abun {int(i)} {float(v)!r}\n")
            env["RHREF_OVERRIDES"] = fov
        _write_input(fin, atmos, np.zeros(1), mu, atm_scale)
        r = subprocess.run([str(exe), mode, str(rundir), fin, fout],
                           env=env, capture_output=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"reference {mode} failed: {r.stderr.decode()[-2000:]}")
        return np.fromfile(fout, np.float64).reshape(ncol, nout, ndep)


def run_ne_from_nh(rundir, atmos, atm_scale=0):
    """pyrh.get_ne_from_nH per column of atmos [ncol, 9, ndep] (rows 0, 1,
    8 used): n_e [ncol, ndep] in cm^-3."""
    return _run_mode("--nefromnh", rundir, atmos, 0.0, atm_scale, 1)[:, 0]


def run_scales(rundir, atmos, lam_ref, atm_scale=0, abun=()):
    """pyrh.get_scales per column of atmos [ncol, 9, ndep]: tau, height [m],
    cmass [kg m^-2] as [ncol, 3, ndep]."""
    return _run_mode("--scales", rundir, atmos, lam_ref, atm_scale, 3, abun=abun)
