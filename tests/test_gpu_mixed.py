"""The mixed-size batch on the GPU: 400 synthetic Fe I Zeeman lines with
Lande factors from the records (LS_LANDE = FALSE) on 2000 wavelengths,
columns of 64 to 256 depth points (pyrh_amd.workloads.mixed_*; run
directory tests/golden/run_mixed).

Against the reference's own output (tests/golden/make_golden_mixed.py,
FP64 inverse), the oracle on seeded batches, and invariants of the sorted
execution: a batch of uneven depth counts runs sorted by depth count, and
every column must come back exactly as it does alone.  Tolerances: I within
1e-10 relative and Q, U, V within 1e-10 I_c for the plain FP64 path; the
FP32 form within 1e-4 relative in I and 1e-5 I_c in Q, U and V, against the
reference's output and against the FP64 path.  At the cores of the
strongest of these 400 lines |q| -> 1: the FP32 form evaluates the Voigt
functions in float but combines them, and runs the DELO recursion, in FP64
(float points there reached 9e-5 I_c)."""
import numpy as np
import pytest

from conftest import GOLDEN
from pyrh_amd import model, workloads as W

pytestmark = pytest.mark.gpu

RUN = GOLDEN / "run_mixed"
TOL = 1e-10


@pytest.fixture(scope="module")
def mixed_fix():
    return np.load(GOLDEN / "mixed_fixtures.npz")


@pytest.fixture(scope="module")
def mdl(mixed_fix):
    return model.build_model(model.open_run(str(RUN)), mixed_fix["wave"])


@pytest.fixture(scope="module")
def eng(mdl):
    from pyrh_amd.engine import Engine
    e = Engine(mdl)
    yield e
    e.close()


def _errs(st, ref):
    Ic = ref[:, 0, :1]
    eI = float(np.max(np.abs(st[:, 0] - ref[:, 0]) / np.abs(ref[:, 0])))
    eP = float(np.max(np.abs(st[:, 1:] - ref[:, 1:]) / Ic[:, None]))
    return eI, eP


@pytest.mark.parametrize("name", ("mixed_mu1", "mixed_mu04"))
def test_mixed_matches_reference(eng, mixed_fix, name):
    atm, nd = mixed_fix[name + "__atmos"], mixed_fix[name + "__ndep"]
    st = eng.synth_lte(atm, mu=float(mixed_fix[name + "__mu"]), ndep=nd)
    eI, eP = _errs(st, mixed_fix[name + "__stokes_o2"])
    print(f"{name}: ndep {nd.tolist()}  I {eI:.2e}  QUV/Ic {eP:.2e}", flush=True)
    assert eI < TOL and eP < TOL


def test_mixed_batch_sorted(eng, mdl):
    """192 columns (sorted execution) against the same columns in a batch
    too small to be sorted, against the oracle, and the FP32 form."""
    import pyoracle
    atm, nd = W.mixed_batch(192, seed=W.SEED_BASE + 31)
    st = eng.synth_lte(atm, mu=0.7, ndep=nd)
    alone = eng.synth_lte(atm[:100], mu=0.7, ndep=nd[:100])
    np.testing.assert_array_equal(st[:100], alone)
    assert np.all(np.isfinite(st))
    sel = np.array([0, 1, 2, 3, 150, 191])
    ref = pyoracle.synth_lte(mdl, atm[sel], mu=0.7, ndep=nd[sel])
    eI, eP = _errs(st[sel], ref)
    print(f"sorted batch vs oracle: I {eI:.2e}  QUV/Ic {eP:.2e}", flush=True)
    assert eI < TOL and eP < TOL
    st32 = eng.synth_lte(atm, mu=0.7, ndep=nd, precision=1)
    Ic = st[:, 0, :1]
    dI = float(np.max(np.abs(st32[:, 0] - st[:, 0]) / Ic))
    dP = float(np.max(np.abs(st32[:, 1:] - st[:, 1:]) / Ic[:, None]))
    print(f"FP32 form: |dI|/Ic {dI:.2e}  |dQUV|/Ic {dP:.2e}", flush=True)
    assert dI < 1e-4 and dP < 1e-5


@pytest.mark.parametrize("name", ("mixed_mu1", "mixed_mu04"))
def test_mixed_fp32_matches_reference(eng, mixed_fix, name):
    """The FP32 form directly against the reference's output: I within 1e-4
    relative, Q, U, V within 1e-5 I_c."""
    atm, nd = mixed_fix[name + "__atmos"], mixed_fix[name + "__ndep"]
    st = eng.synth_lte(atm, mu=float(mixed_fix[name + "__mu"]), ndep=nd, precision=1)
    eI, eP = _errs(st, mixed_fix[name + "__stokes_o2"])
    print(f"{name} FP32 form: I {eI:.2e}  QUV/Ic {eP:.2e}", flush=True)
    assert eI < 1e-4 and eP < 1e-5


def test_mixed_device_pointers(eng):
    import torch
    atm, nd = W.mixed_batch(160, seed=W.SEED_BASE + 32)
    host = eng.synth_lte(atm, mu=0.4, ndep=nd, precision=1)
    dev = torch.device("cuda:0")
    d_atm = torch.from_numpy(atm).to(dev)
    d_nd = torch.from_numpy(nd).to(dev)
    d_out = torch.empty((atm.shape[0], 4, eng.nout), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    eng.synth_lte_device(d_atm.data_ptr(), d_nd.data_ptr(), atm.shape[0], atm.shape[2],
                         d_out.data_ptr(), mu=0.4, precision=1)
    np.testing.assert_array_equal(d_out.cpu().numpy(), host)


def _synth_device(eng, atm, nd, mu=0.7, precision=1):
    """rhg_synth_lte on device buffers; the output starts as NaN so that a
    column the kernels never write shows."""
    import torch
    dev = torch.device("cuda:0")
    d_atm = torch.from_numpy(np.ascontiguousarray(atm)).to(dev)
    d_nd = torch.from_numpy(np.ascontiguousarray(nd)).to(dev)
    out = torch.full((atm.shape[0], 4, eng.nout), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    eng.synth_lte_device(d_atm.data_ptr(), d_nd.data_ptr(), atm.shape[0], atm.shape[2],
                         out.data_ptr(), mu=mu, precision=precision)
    return out.cpu().numpy()


def test_device_sort_many_chunks(eng):
    """2500 uneven columns on device buffers (three workgroups of the
    on-device counting sort by depth count): with sorting on, every column
    bit for bit as with sorting off and as in the reversed batch, and every
    output written."""
    atm, nd = W.mixed_batch(2500, seed=W.SEED_BASE + 34)
    try:
        s1 = _synth_device(eng, atm, nd)
        eng.set_tuning("column_sort", 0)
        s0 = _synth_device(eng, atm, nd)
    finally:
        eng.set_tuning("column_sort", 1)
    sr = _synth_device(eng, atm[::-1], nd[::-1])
    assert np.all(np.isfinite(s1))
    np.testing.assert_array_equal(s1, s0)
    np.testing.assert_array_equal(s1, sr[::-1])


@pytest.mark.parametrize("sort", (1, 0))
def test_device_ndep_out_of_range(eng, sort):
    """A depth count outside [4, ndep_max] in device memory: the call
    returns the error (the kernels ran on clamped counts, inside the
    buffers), and the next good call on the context is unaffected."""
    from pyrh_amd.engine import EngineError
    atm, nd = W.mixed_batch(300, seed=W.SEED_BASE + 35)
    good = _synth_device(eng, atm, nd)
    try:
        eng.set_tuning("column_sort", sort)
        for bad in (3, atm.shape[2] + 1, -7):
            nd2 = nd.copy()
            nd2[137] = bad
            with pytest.raises(EngineError, match="ndep out of range"):
                _synth_device(eng, atm, nd2)
    finally:
        eng.set_tuning("column_sort", 1)
    np.testing.assert_array_equal(_synth_device(eng, atm, nd), good)


def test_mixed_rf_sorted(eng):
    """Response functions of a sorted batch: each column as it comes alone,
    zero past its depth, and the T response against central differences
    of the FP64 synthesis on one column."""
    atm, nd = W.mixed_batch(130, seed=W.SEED_BASE + 33)
    st, rf = eng.rf_lte(atm, mu=0.7, ndep=nd, precision=0)
    st8, rf8 = eng.rf_lte(atm[:8], mu=0.7, ndep=nd[:8], precision=0)
    np.testing.assert_array_equal(st[:8], st8)
    np.testing.assert_array_equal(rf[:8], rf8)
    for c in range(0, 130, 13):
        assert np.all(rf[c, :, nd[c]:] == 0.0)
    assert np.all(np.isfinite(rf))
    # d I / d T at a few depths of column 0 against central differences
    c, n = 0, int(nd[0])
    one = atm[c:c + 1]
    for k in (n // 4, n // 2, 3 * n // 4):
        p, m = one.copy(), one.copy()
        p[0, 1, k] += 0.1
        m[0, 1, k] -= 0.1
        fd = (eng.synth_lte(p, 0.7, ndep=nd[c:c + 1]) - eng.synth_lte(m, 0.7, ndep=nd[c:c + 1]))[0] / 0.2
        ad = rf[c, 0, k]
        scale = np.max(np.abs(fd[0]))
        assert scale > 0
        err = np.max(np.abs(ad[0] - fd[0])) / scale
        print(f"dI/dT depth {k}: max err {err:.2e} of max |dI/dT|", flush=True)
        assert err < 1e-3
