"""Hydrostatic equilibrium on the GPU (rhg_hse, pyrh_amd.api.hse) against
the reference's own hse() (golden vectors from the reference library
compiled from its C sources, tests/golden/make_golden_hse.py) and against
the oracle restatement on seeded batches with ragged depth counts.

Tolerance: n_e, n_Htot, rho, p_g within 1e-10 relative of the reference
and of the oracle.  The GPU follows the reference's operation order; its
exp/pow/log10 round differently from glibc in the last place, and the
lanes that share a column sum the elements of the electron-density step
in a different order, so results agree to round-off, not bit for bit."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from test_oracle_hse import CASES, hse_model, relerr

pytestmark = pytest.mark.gpu

TOL = 1e-10
RUN = GOLDEN / "run_falc"


@pytest.fixture(scope="module")
def fix():
    return np.load(GOLDEN / "hse_fixtures.npz")


_ENG = {}


def engine(fix, name):
    from pyrh_amd.engine import Engine
    key = name if name in ("fudge", "abund") else "base"
    if key not in _ENG:
        hm = hse_model(fix, key if key != "base" else "single")
        _ENG[key] = (Engine(hm.model), hm)
    return _ENG[key]


def _run(fix, name, lanes=None):
    e, hm = engine(fix, name)
    old = os.environ.pop("RHG_HSE_LANES", None)
    if lanes:
        os.environ["RHG_HSE_LANES"] = str(lanes)
    try:
        out = e.hse(hm, fix[name + "__scale"], fix[name + "__T"], float(fix[name + "__pg_top"]),
                    int(fix[name + "__atm_scale"]), fix[name + "__ndep"])
        return out, e.hse_timing()[1]
    finally:
        os.environ.pop("RHG_HSE_LANES", None)
        if old is not None:
            os.environ["RHG_HSE_LANES"] = old


@pytest.mark.parametrize("name", CASES)
def test_hse_matches_reference(fix, name):
    nd = fix[name + "__ndep"]
    ref = fix[name + "__out"]
    for lanes in (None, 1):
        (ne, nH, rho, pg, it), lpc = _run(fix, name, lanes)
        worst = 0.0
        for c in range(nd.size):
            n = nd[c]
            got = np.stack([ne[c], nH[c], rho[c], pg[c]])
            assert np.all(got[:, n:] == 0.0) and np.all(it[c, n:] == 0)
            worst = max(worst, relerr(got[:, :n], ref[c, :, :n]))
        print(f"{name}: lanes/column {lpc}  max rel err {worst:.2e}", flush=True)
        assert worst < TOL


def test_hse_batch_matches_oracle(fix):
    """96 seeded map columns, ragged, against the oracle; the iteration
    counts per layer agree as well."""
    import pyoracle
    from pyrh_amd import workloads as W
    e, hm = engine(fix, "single")
    T = W.stokes_map(96, seed=W.SEED_BASE + 42)[:, 1]
    nd = np.full(96, W.NDEP, np.int32)
    nd[7], nd[30], nd[61] = W.NDEP - 1, 50, 12
    ne, nH, rho, pg, it = e.hse(hm, W.LOGTAU, T, 0.1, 0, nd)
    one, two, three, four, oit = pyoracle.hse(hm, np.broadcast_to(W.LOGTAU, T.shape), T, 0.1,
                                              ndep=nd)
    err = max(relerr(a[c, :nd[c]], b[c, :nd[c]]) for a, b in ((ne, one), (nH, two), (rho, three),
                                                              (pg, four)) for c in range(96))
    print(f"96 columns vs oracle: max rel err {err:.2e}  lanes {e.hse_timing()[1]}", flush=True)
    assert err < TOL
    np.testing.assert_array_equal(it, oit)
    # a column does not depend on its neighbours
    perm = np.random.default_rng(5).permutation(96)
    p2 = e.hse(hm, W.LOGTAU, T[perm], 0.1, 0, nd[perm])
    np.testing.assert_array_equal(p2[3], pg[perm])
    np.testing.assert_array_equal(p2[0], ne[perm])


def test_hse_device_pointers(fix):
    import torch
    from pyrh_amd import workloads as W
    e, hm = engine(fix, "single")
    T = W.stokes_map(40, seed=W.SEED_BASE + 43)[:, 1]
    sc = np.ascontiguousarray(np.broadcast_to(W.LOGTAU, T.shape))
    nd = np.full(40, W.NDEP, np.int32)
    nd[3] = 33
    pgt = np.linspace(0.05, 0.2, 40)
    host = e.hse(hm, sc, T, pgt, 0, nd)
    dev = torch.device("cuda:0")
    d = {k: torch.from_numpy(v).to(dev) for k, v in
         (("sc", sc), ("T", np.ascontiguousarray(T)), ("pg", pgt), ("nd", nd))}
    outs = [torch.empty(T.shape, dtype=torch.float64, device=dev) for _ in range(4)]
    it = torch.empty(T.shape, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    e.hse_device(hm, d["sc"].data_ptr(), d["T"].data_ptr(), d["pg"].data_ptr(),
                 d["nd"].data_ptr(), 40, W.NDEP, *(o.data_ptr() for o in outs), 0,
                 it.data_ptr())
    for a, b in zip(outs, host[:4]):
        np.testing.assert_array_equal(a.cpu().numpy(), b)
    np.testing.assert_array_equal(it.cpu().numpy(), host[4])


def test_api_hse(fix):
    """pyrh.hse's signature and returns; the workload structure
    (pyrh_amd/data/lte_atmos82.npz, made with pyrh's hse()) comes back."""
    from pyrh_amd import api, workloads as W
    d = np.load(W.DATA / "lte_atmos82.npz")
    ne, nH = api.hse(str(GOLDEN / "run_lte"), 0, d["logtau"], d["T"], pg_top=0.1)
    assert relerr(ne * 1e-6, d["ne_cm3"]) < TOL and relerr(nH * 1e-6, d["nH_cm3"]) < TOL
    ne4, nH4, rho, pg = api.hse(str(GOLDEN / "run_lte"), 0, d["logtau"], d["T"], full_output=True)
    np.testing.assert_array_equal(ne4, ne)
    assert relerr(pg, d["pg_si"]) < TOL and relerr(rho, d["rho_si"]) < TOL
    name = "fudge"
    out = api.hse(str(RUN), 0, fix[name + "__scale"][0], fix[name + "__T"][0],
                  fudge_wave=fix["fudge_wave"], fudge_value=fix["fudge_value"], full_output=True)
    assert relerr(np.stack(out), fix[name + "__out"][0]) < TOL
    name = "abund"
    out = api.hse(str(RUN), 0, fix[name + "__scale"][0], fix[name + "__T"][0],
                  atomic_number=fix["abun_id"], atomic_abundance=fix["abun_val"], full_output=True)
    assert relerr(np.stack(out), fix[name + "__out"][0]) < TOL
    b = api.hse_batch(str(RUN), 0, fix["map6__scale"], fix["map6__T"])
    assert relerr(np.stack([b["ne"], b["nHtot"], b["rho"], b["pg"]], 1), fix["map6__out"]) < TOL


def test_hse_large_batch_invariants():
    """4096 columns: the ideal-gas law of the final layer state, p_g rising
    with depth, finite and positive densities."""
    from pyrh_amd import api, workloads as W
    T = W.stokes_map(4096, seed=W.SEED_BASE + 44)[:, 1]
    b = api.hse_batch(str(GOLDEN / "run_lte"), 0, W.LOGTAU, T, 0.1)
    e, hm = api.hse_engine_for(str(GOLDEN / "run_lte"))
    s = hm.model.scalars
    ideal = (b["nHtot"] * s["totalAbund"] + b["ne"]) * 1.380658E-23 * T
    np.testing.assert_allclose(ideal, b["pg"], rtol=1e-12)
    assert np.all(np.diff(b["pg"], axis=1) > 0.0)
    assert np.all(b["ne"] > 0.0) and np.all(np.isfinite(b["rho"]))
    assert np.all((b["niter"] >= 1) & (b["niter"] <= hm.nmax_iter))


@pytest.mark.parametrize("name", ("ne_map4", "ne_hot", "ne_falc"))
def test_ne_from_nh_matches_reference(fix, name):
    from pyrh_amd import api
    a = fix[name + "__atmos"]
    ref = fix[name + "__ne"]
    e, hm = api.hse_engine_for(str(RUN))
    ne = e.ne_from_nh(hm, a[:, 1], a[:, 8])
    err = relerr(ne, ref)
    one = api.get_ne_from_nH(str(RUN), 0, a[0, 0], a[0, 1], a[0, 8])
    print(f"{name}: max rel err {err:.2e}", flush=True)
    assert err < TOL
    np.testing.assert_array_equal(one, ne[0])


@pytest.mark.parametrize("name", ("sc_tau", "sc_tau400", "sc_abund", "sc_cmass", "sc_height"))
def test_scales_match_reference(fix, name):
    from pyrh_amd import api
    from test_oracle_hse import SC_FILLED
    a = fix[name + "__atmos"]
    lam = float(fix[name + "__lam_ref"])
    ats = int(fix[name + "__atm_scale"])
    kw = {}
    if name == "sc_abund":
        kw = dict(atomic_number=fix["abun_id"], atomic_abundance=fix["abun_val"])
    e = api.scales_engine_for(str(RUN), lam, **kw)
    got = np.stack(e.scales(a, ats), 1)
    ref = fix[name + "__scales"]
    worst = 0.0
    for i in SC_FILLED[ats]:
        scale = np.max(np.abs(ref[:, i]), axis=1, keepdims=True)
        worst = max(worst, float(np.max(np.abs(got[:, i] - ref[:, i]) / scale)))
    print(f"{name}: max err {worst:.2e} of the column's range", flush=True)
    assert worst < TOL
    # the input scale comes back in SI units
    if ats == 2:
        own = a[:, 0] * 1e3
    else:
        own = 10.0 ** a[:, 0] * (10.0 if ats == 1 else 1.0)
    np.testing.assert_allclose(got[:, {0: 0, 1: 2, 2: 1}[ats]], own, rtol=1e-13)
    t1, h1, m1 = api.get_scales(str(RUN), ats, a[0, 0], a[0], lam, **kw)
    np.testing.assert_array_equal(np.stack([t1, h1, m1]), got[0])
