"""The launch-configuration knobs (rhg_set_tuning) and the autotuner: every
setting gives bit-identical Stokes vectors (the tuner checks it on the
batch it times), the chosen setting is left on the context, and unknown
keys and values are rejected."""
import numpy as np
import pytest

from pyrh_amd import model, workloads as W
from pyrh_amd.engine import EngineError, TUNE_KEYS

pytestmark = pytest.mark.gpu


def test_autotune_map_and_mixed(setups):
    import torch
    from pyrh_amd import autotune
    from pyrh_amd.engine import Engine
    dev = torch.device("cuda:0")
    eng = Engine(model.build_model(setups["lte"], W.wavelengths_630()))
    atm = torch.from_numpy(W.stokes_map(512, seed=W.SEED_BASE + 31)).to(dev)
    nd = torch.full((512,), atm.shape[2], dtype=torch.int32, device=dev)
    out = torch.empty((512, 4, eng.nout), dtype=torch.float64, device=dev)
    for prec in (1, 0):
        rep = autotune.tune(eng, atm, nd, out, precision=prec, reps=1)
        print(prec, rep["chosen"], [(c["line_table"], c["column_sort"], c["wave_order"], c["ms"])
                                    for c in rep["candidates"]], flush=True)
        assert eng.tuning() == rep["chosen"]
        assert len(rep["candidates"]) == 2 ** len(TUNE_KEYS)
    eng.close()
    # a batch of uneven depth counts, where the column sort matters
    from conftest import GOLDEN
    mdl = model.build_model(model.open_run(str(GOLDEN / "run_mixed")), W.mixed_wavelengths())
    em = Engine(mdl)
    a, n = W.mixed_batch(256, seed=W.SEED_BASE + 32)
    atm = torch.from_numpy(a).to(dev)
    nd = torch.from_numpy(n).to(dev)
    out = torch.empty((256, 4, em.nout), dtype=torch.float64, device=dev)
    rep = autotune.tune(em, atm, nd, out, mu=0.7, precision=1, reps=1)
    print("mixed", rep["chosen"], flush=True)
    em.close()


def test_tuning_rejects_unknown(setups):
    import ctypes as C
    from pyrh_amd.engine import Engine, lib
    eng = Engine(model.build_model(setups["lte"], W.wavelengths_630()))
    with pytest.raises(EngineError):
        eng.set_tuning("line_table", 2)
    assert lib().rhg_set_tuning(eng._h, 7, 0) != 0
    v = C.c_int32()
    assert lib().rhg_get_tuning(eng._h, 7, C.byref(v)) != 0
    eng.set_tuning("wave_order", 1)
    assert eng.tuning()["wave_order"] == 1
    eng.set_tuning("wave_order", 0)
    eng.close()
