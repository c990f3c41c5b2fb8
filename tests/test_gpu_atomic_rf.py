"""d I / d log gf on the GPU, the reference's analytic atomic response
functions, against the reference's own output for the same inputs (golden
cases written by rhf1d built from the reference sources, FP64 matrix
inverse): a NO_STOKES run with two log gf parameters and one wavelength-shift
parameter, and a FULL_STOKES run with overridden log gf values."""
import numpy as np
import pytest

from conftest import RUNS, golden_model, override_kwargs
from pyrh_amd import model, workloads

pytestmark = pytest.mark.gpu

CASES = ("rf_nostokes", "overrides")
TOL = 1e-9       # relative to the largest |d I / d log gf| of the column and parameter


def _case(fixtures, setups, name):
    from pyrh_amd.engine import Engine
    m = golden_model(fixtures, setups, name)
    kw = override_kwargs(fixtures, name)
    par = model.atomic_rf_lines(m.setup, kw.get("loggf_ids"), len(kw.get("lam_ids", [])))
    return Engine(m), m, kw, par


@pytest.mark.parametrize("name", CASES)
def test_loggf_rf_matches_reference(fixtures, setups, name):
    eng, m, kw, par = _case(fixtures, setups, name)
    atm = fixtures[name + "__atmos"]
    mu, scale = float(fixtures[name + "__mu"]), int(fixtures[name + "__scale"])
    st, rf = eng.rf_loggf(atm, par, mu, scale)
    ref = fixtures[name + "__rf_loggf"]               # [ncol, nlw, nloggf]
    ng = ref.shape[2]
    big = np.max(np.abs(ref), axis=1, keepdims=True)
    if name == "rf_nostokes":
        assert np.all(big > 0), "every log gf parameter responds somewhere"
    # all-zero columns (FULL_STOKES: the polarized solver fills no response)
    # must come back exactly zero
    err = np.abs(rf[:, :, :ng] - ref) / np.where(big > 0, big, 1e-300)
    print(f"{name}: npar {par.size}  max err {err.max():.2e}", flush=True)
    assert err.max() < TOL
    # the wavelength-shift parameters are never filled by the reference
    assert np.all(rf[:, :, ng:] == 0.0)
    # zero exactly where the reference's solver leaves it zero
    assert np.array_equal(rf[:, :, :ng] == 0.0, ref == 0.0)
    np.testing.assert_array_equal(st, eng.synth_lte(atm, mu, scale))


def test_loggf_rf_matches_oracle_on_map_columns(setups):
    """64 seeded map columns of the NO_STOKES run, both Fe I lines as
    parameters, against the oracle restatement."""
    import pyoracle
    from pyrh_amd.engine import Engine
    wave = workloads.wavelengths_630()
    s = model.open_run(str(RUNS["nostokes"]), loggf_ids=np.array([0, 1], np.int32),
                       loggf_values=np.array([-0.71, -0.969]))
    m = model.build_model(s, wave)
    par = model.atomic_rf_lines(s, [0, 1], 0)
    atm = workloads.stokes_map(64, seed=workloads.SEED_BASE + 9)
    eng = Engine(m)
    st, rf = eng.rf_loggf(atm, par)
    cst, crf = pyoracle.rf_loggf(m, atm, par)
    big = np.max(np.abs(crf), axis=1, keepdims=True)
    assert np.all(big > 0)
    err = np.abs(rf - crf) / big
    print(f"map columns: max err {err.max():.2e}", flush=True)
    assert err.max() < TOL
    assert np.max(np.abs(st[:, 0] - cst[:, 0]) / cst[:, 0]) < 1e-10


def test_compute1d_get_atomic_rfs(fixtures):
    from pyrh_amd import api
    name = "rf_nostokes"
    kw = override_kwargs(fixtures, name)
    atm = fixtures[name + "__atmos"][0]
    (I, Q, U, V, lam), rf = api.compute1d(str(RUNS[str(fixtures[name + "__run"])]),
                                          float(fixtures[name + "__mu"]),
                                          int(fixtures[name + "__scale"]), atm,
                                          fixtures[name + "__wave"], get_atomic_rfs=True, **kw)
    ref = fixtures[name + "__rf_loggf"][0]            # [nlw, nloggf]
    assert rf.shape == (len(kw["loggf_ids"]) + len(kw["lam_ids"]), lam.size)
    big = np.max(np.abs(ref), axis=0)
    assert np.max(np.abs(rf[:ref.shape[1]].T - ref) / big) < TOL
    np.testing.assert_allclose(I, fixtures[name + "__stokes_o2"][0, 0], rtol=1e-10, atol=0)


@pytest.mark.parametrize("tag", ("mu1", "mu07"))
def test_loggf_rf_matches_reference_itop(tag):
    """d I / d log gf with ATMOS_ITOP (NO_STOKES, moving columns): the inward
    ray starts at the irradiating intensity (rh/rhf1d/bezier_1D.c:378-379),
    GPU against the reference's own output."""
    from conftest import GOLDEN
    from pyrh_amd.engine import Engine
    f = np.load(GOLDEN / "static_fixtures.npz")
    lg = f["itop_loggf"]
    kw = dict(loggf_ids=lg[:, 0].astype(np.int32), loggf_values=lg[:, 1])
    s = model.open_run(str(GOLDEN / "run_lte_itop_nostokes"), **kw)
    m = model.build_model(s, f["itop_wave"])
    par = model.atomic_rf_lines(s, kw["loggf_ids"], 0)
    eng = Engine(m)
    st, rf = eng.rf_loggf(f["itop_atmos"], par, float(f[f"itop_rf_{tag}__mu"]), 0)
    ref, sref = f[f"itop_rf_{tag}__rf"], f[f"itop_rf_{tag}__stokes"]
    err = float(np.max(np.abs(rf - ref)) / np.max(np.abs(ref)))
    eI = float(np.max(np.abs(st[:, 0] - sref[:, 0]) / sref[:, 0]))
    print(f"itop rf {tag}: rf {err:.2e}  I {eI:.2e}", flush=True)
    assert err < TOL and eI < 1e-10
