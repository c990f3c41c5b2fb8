import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "oracle", ROOT / "tests" / "golden"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))
os.environ.setdefault("PYRH_PATH", str(ROOT / "tests" / "golden" / "rhdata"))

GOLDEN = ROOT / "tests" / "golden"
RUNS = {"lte": GOLDEN / "run_lte", "nostokes": GOLDEN / "run_lte_nostokes",
        "unpol": GOLDEN / "run_lte_unpol", "falc": GOLDEN / "run_falc"}


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def fixtures():
    import numpy as np
    return np.load(GOLDEN / "lte_fixtures.npz")


@pytest.fixture(scope="session")
def setups():
    from pyrh_amd import model
    return {k: model.open_run(str(v)) for k, v in RUNS.items()}


def fixture_cases(F):
    return sorted({k.split("__")[0] for k in F.files})


def override_kwargs(F, name):
    """pyrh.compute1d keyword overrides stored with a golden case."""
    kw = {}
    if name + "__ov_loggf" not in F.files:
        return kw
    lg, lm, ab = (F[f"{name}__ov_{k}"] for k in ("loggf", "lam", "abun"))
    if len(lg):
        kw["loggf_ids"] = lg[:, 0].astype(np.int32)
        kw["loggf_values"] = lg[:, 1]
    if len(lm):
        kw["lam_ids"] = lm[:, 0].astype(np.int32)
        kw["lam_values"] = lm[:, 1]
    if len(ab):
        kw["atomic_number"] = ab[:, 0].astype(np.int32)
        kw["atomic_abundance"] = ab[:, 1]
    return kw


_MODELS = {}


def golden_model(F, setups, name):
    """The host model of a golden case (run directory, wavelength grid and
    any keyword overrides), cached."""
    from pyrh_amd import model
    if name not in _MODELS:
        kw = override_kwargs(F, name)
        run = str(F[name + "__run"])
        s = model.open_run(str(RUNS[run]), **kw) if kw else setups[run]
        _MODELS[name] = model.build_model(s, F[name + "__wave"])
    return _MODELS[name]
