"""Host logic: input readers, model builder and the C ABI library (no GPU)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from pyrh_amd import abi, constants as K, model, workloads
from pyrh_amd.cscan import sscanf
from pyrh_amd.rhdata import RHInputError, read_keywords

from conftest import GOLDEN

ROOT = Path(__file__).resolve().parents[1]


def test_sscanf_semantics():
    assert sscanf("630.1500-10.123 26.00", "%lf %lf %s") == (3, [630.15, -10.123, "26.00"])
    assert sscanf("26.01", "%d.%d") == (2, [26, 1])
    assert sscanf("   \n", "%lf") == (-1, [])
    assert sscanf(" 1503 1835", "%5d%5d") == (2, [1503, 1835])
    assert sscanf("5D", "%d%1s") == (2, [5, "D"])


def test_keyword_defaults_and_required(tmp_path):
    kw = read_keywords(ROOT / "tests/golden/run_lte/keyword.input")
    assert kw.str("STOKES_MODE") == "FULL_STOKES"
    assert kw.float("LAMBDA_REF") == 500.0        # default
    assert not kw.bool("MAGNETO_OPTICAL")
    p = tmp_path / "keyword.input"
    p.write_text("  NRAYS = 1\n")
    with pytest.raises(RHInputError):
        read_keywords(p)


def test_air_to_vacuum_known_value():
    # 630.1500 nm (air) -> 630.3245 nm (vacuum) for the Fe I line
    assert abs(K.air_to_vacuum(630.15) - 630.32427) < 1e-4
    # the two directions are separate fits (Edlen 1966 and its inverse
    # series), consistent to about 1e-6 relative
    lam = np.array([500.0, 630.0, 1000.0])
    back = np.array(K.vacuum_to_air(K.air_to_vacuum(lam)))
    np.testing.assert_allclose(back, lam, rtol=2e-6)
    assert K.air_to_vacuum(150.0) == 150.0        # below 200 nm nothing changes


def test_fe_line_zeeman_patterns(setups):
    """Fe I 630.15 (z5P2-e5D2: 13 components, g_eff 1.667) and 630.25
    (z5P1-e5D0: 3 components, g_eff 2.5); strengths sum to 1 per q."""
    rlk = setups["lte"].rlk
    assert [len(r.comps) for r in rlk] == [13, 3]
    for r, geff in zip(rlk, (5.0 / 3.0, 2.5)):
        for q in (-1, 0, 1):
            s = sum(c[2] for c in r.comps if c[0] == q)
            if any(c[0] == q and c[2] > 0 for c in r.comps):
                assert abs(s - 1) < 1e-14
        g = sum(c[1] * c[2] for c in r.comps if c[0] == 1)
        assert abs(g - geff) < 1e-12
        assert r.vdwaals == 2              # ABO (Anstee-Barklem-O'Mara) broadening
        assert r.polarizable


def test_unpolarizable_labels(setups):
    assert not any(r.polarizable for r in setups["unpol"].rlk)


def test_model_solvers_and_tables(setups):
    m = model.build_model(setups["lte"], workloads.wavelengths_630())
    assert m.scalars["nspect"] == 113 and m.lam[m.ref_index] == 500.0
    assert np.all(m.arrays["solver"][m.out_index] == 2)
    assert m.arrays["solver"][m.ref_index] == 0
    # every output wavelength sees both lines (window 20 vmicro_char)
    ptr = m.arrays["lam_line_ptr"]
    assert np.all(np.diff(ptr)[m.out_index] == 2)
    # continua active at 500 nm and 630 nm exist and have positive cross-sections
    assert np.all(m.arrays["lam_cont_alpha"] > 0)


def test_model_rejects_unsupported(setups, tmp_path):
    import shutil
    d = tmp_path / "run"
    shutil.copytree(ROOT / "tests/golden/run_lte", d)
    kw = (d / "keyword.input").read_text().replace("RLK_SCATTER = FALSE",
                                                   "RLK_SCATTER = TRUE")
    (d / "keyword.input").write_text(kw)
    s = model.open_run(str(d))
    with pytest.raises(NotImplementedError):
        model.build_model(s, workloads.wavelengths_630())


def test_model_magneto_optical_switch():
    """MAGNETO_OPTICAL = TRUE reaches the model's switch (the polarized
    solver then carries the Faraday-Voigt terms)."""
    for run, flag in (("run_lte", 0), ("run_lte_mo", 1)):
        s = model.open_run(str(ROOT / "tests/golden" / run))
        m = model.build_model(s, workloads.wavelengths_630())
        assert m.scalars["magneto_optical"] == flag


def test_workloads_are_seeded():
    a = workloads.stokes_map(8, seed=3)
    b = workloads.stokes_map(8, seed=3)
    np.testing.assert_array_equal(a, b)
    assert a.shape == (8, 9, 82)
    assert np.all(a[:, 5] >= 0) and np.all(a[:, 5] <= 3000)
    assert np.all(np.abs(np.cos(a[:, 6])) <= 1)


def test_abi_struct_layout_matches_oracle():
    import pyoracle
    ms, cs = pyoracle.sizes()
    assert ms == C.sizeof(abi.RhgModel) and cs == C.sizeof(abi.RhgColumns)


def _header_functions():
    txt = (ROOT / "include" / "pyrh_amd.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rhg_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_function():
    """libpyrh_amd.so loads on the host and exports what the header declares."""
    from pyrh_amd import engine
    lib = engine.lib()
    names = _header_functions()
    assert "rhg_synth_lte" in names and "rhg_open" in names
    for n in names:
        assert hasattr(lib, n), n
    assert lib.rhg_abi_version() == abi.ABI_VERSION
    assert lib.rhg_model_sizeof() == C.sizeof(abi.RhgModel)


def test_library_reports_errors_without_device():
    """A null model is rejected with a status code and a message."""
    from pyrh_amd import engine
    lib = engine.lib()
    h = C.c_void_p()
    rc = lib.rhg_open(None, 0, C.byref(h))
    assert rc == -1 and b"null" in lib.rhg_last_error()


def test_hse_host_tables():
    """The HSE model: packed element tables with hydrogen first, the opacity
    wavelength (the last of {LAMBDA_REF, 500 nm}), the fudge factors there
    (Linear over the fudge table) and the rejected column-mass scale."""
    import pytest
    from pyrh_amd import api, hse, model
    s = model.open_run(str(GOLDEN / "run_falc"))
    hm = hse.build_hse(s)
    npf = s.elements.Tpf.size
    assert hm.el_pf_ptr[0] == 0 and hm.el_abund[0] == 1.0
    np.testing.assert_array_equal(np.diff(hm.el_pf_ptr), hm.el_nstage[:-1] * npf)
    assert hm.el_logpf.size == hm.el_nstage.sum() * npf
    assert hm.model.lam[hm.lam_index] == 500.0 and hm.lam_index == hm.model.lam.size - 1
    assert hm.hmin_fudge == 1.0 and hm.scatt_fudge == 1.0
    fw = np.array([300.0, 450.0, 600.0])
    fv = np.array([[1.3, 1.2, 1.1], [0.7, 0.8, 0.9], [1.0, 1.0, 1.0]])
    hf = hse.build_hse(s, fw, fv)
    assert abs(hf.hmin_fudge - (1.2 + (1.1 - 1.2) * 50.0 / 150.0)) < 1e-15
    assert abs(hf.scatt_fudge - (0.8 + (0.9 - 0.8) * 50.0 / 150.0)) < 1e-15
    with pytest.raises(NotImplementedError):
        api.hse(str(GOLDEN / "run_falc"), 1, np.zeros(5), np.full(5, 5000.0))
    with pytest.raises(ValueError):
        hse.build_hse(s, fw, None)


class _FakeTuneEngine:
    """Stands in for Engine in the autotuner's host logic: a synthesis whose
    time depends on the launch configuration and whose output is the same
    for every configuration (or differs for one, to exercise the check)."""

    def __init__(self, torch, out, bad=None):
        self.cfg = {"line_table": 1, "column_sort": 1, "wave_order": 0}
        self.torch, self.out, self.bad, self.calls = torch, out, bad, []

    def set_tuning(self, key, value):
        self.cfg[key] = int(value)

    def tuning(self):
        return dict(self.cfg)

    def timings(self):
        return {"formal": 1.0}

    def synth_lte_device(self, *args, **kw):
        import time
        self.calls.append(dict(self.cfg))
        # line table on is fastest, column sort second
        time.sleep(0.002 * (2 - self.cfg["line_table"]) + 0.001 * (1 - self.cfg["column_sort"]))
        self.out.fill_(1.0)
        if self.bad is not None and all(self.cfg[k] == v for k, v in self.bad.items()):
            self.out[0, 0, 0] = 2.0


def test_autotune_picks_fastest_and_checks_bit_identity(monkeypatch):
    """pyrh_amd.autotune.tune (host logic, no GPU): every combination of the
    launch knobs is timed, the fastest is left on the engine, and a
    configuration that changes the Stokes vectors is an error."""
    torch = pytest.importorskip("torch")
    from pyrh_amd import autotune
    monkeypatch.setattr(torch.cuda, "synchronize", lambda *a, **k: None)
    atm = torch.zeros((4, 9, 8), dtype=torch.float64)
    ndep = torch.full((4,), 8, dtype=torch.int32)
    out = torch.zeros((4, 4, 3), dtype=torch.float64)
    eng = _FakeTuneEngine(torch, out)
    rep = autotune.tune(eng, atm, ndep, out, reps=2)
    assert len(rep["candidates"]) == 8
    assert {(c["line_table"], c["column_sort"], c["wave_order"]) for c in rep["candidates"]} == \
        {(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)}
    assert rep["chosen"]["line_table"] == 1 and rep["chosen"]["column_sort"] == 1
    assert eng.tuning() == rep["chosen"]
    bad = _FakeTuneEngine(torch, out, bad={"line_table": 0, "column_sort": 1, "wave_order": 0})
    with pytest.raises(RuntimeError, match="changed the Stokes vectors"):
        autotune.tune(bad, atm, ndep, out, reps=1)


def test_grid_positions_and_table():
    """The wavelength grids of one merged context: each grid's outputs are
    located on the merged model's output wavelengths (what compute1d with
    that wave returns: sorted, duplicates and LAMBDA_REF dropped), and the
    (grid_ptr, grid_out) table of rhg_set_grids concatenates them."""
    from pyrh_amd import api
    from pyrh_amd.engine import grid_table
    run = model.open_run(str(GOLDEN / "run_lte"))
    wl = workloads.wavelengths_630()
    lam_ref = run.kw.float("LAMBDA_REF")
    waves = [wl[::-1][:40], wl[30:], np.concatenate([wl[::9], wl[::9], [lam_ref]]), wl[:0]]
    merged = np.unique(np.concatenate(waves))
    m = model.build_model(run, merged)
    pos = api.grid_positions(m, waves)
    out_lam = m.lam[m.out_index]
    for w, p in zip(waves, pos):
        single = model.build_model(run, w) if w.size else None
        want = single.lam[single.out_index] if single is not None else np.zeros(0)
        np.testing.assert_array_equal(out_lam[p], want)
        assert np.all(np.diff(p) > 0)
    ptr, flat = grid_table(pos)
    assert ptr.tolist() == [0, 40, 40 + 82, 40 + 82 + pos[2].size, flat.size]
    np.testing.assert_array_equal(flat[ptr[1]:ptr[2]], pos[1])
    with pytest.raises(ValueError, match="not on the merged grid"):
        api.grid_positions(m, [np.array([wl[0] + 1e-4])])
