import os
import sys
from pathlib import Path

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
