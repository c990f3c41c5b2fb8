"""rhg_synth_lte_noalloc (the synthesis without device allocation, its work
buffers in a fixed device array) through the C ABI, against rhg_synth_lte on
the same device inputs: bit-identical Stokes vectors on two calls in a row,
the second smaller, in both precision forms; an oversized batch and an
out-of-range depth count are rejected."""
import ctypes as C

import numpy as np
import pytest

from pyrh_amd import model, workloads as W
from pyrh_amd.abi import RhgColumns

pytestmark = pytest.mark.gpu

RHG_EINVAL = -1


def _cols(d_atm, d_nd, ncol, nd, mu=1.0):
    s = RhgColumns()
    s.ncol, s.ndep_max = int(ncol), int(nd)
    s.ndep = C.cast(C.c_void_p(d_nd.data_ptr()), C.POINTER(C.c_int32))
    s.atmos = C.cast(C.c_void_p(d_atm.data_ptr()), C.POINTER(C.c_double))
    s.atm_scale, s.mu = 0, float(mu)
    return s


@pytest.fixture(scope="module")
def eng(setups):
    from pyrh_amd.engine import Engine
    e = Engine(model.build_model(setups["lte"], W.wavelengths_630()))
    yield e
    e.close()


@pytest.mark.parametrize("precision,ncols", [(1, (600, 150)), (0, (480, 100))])
def test_noalloc_matches_synth(eng, precision, ncols):
    import torch
    from pyrh_amd.engine import lib
    L = lib()
    dev = torch.device("cuda:0")
    for call, ncol in enumerate(ncols):
        atm = W.stokes_map(ncol, seed=W.SEED_BASE + 90 + call)
        nd = np.full(ncol, 82, np.int32)
        nd[::7] = 70
        d_atm = torch.from_numpy(atm).to(dev)
        d_nd = torch.from_numpy(nd).to(dev)
        ref = torch.empty((ncol, 4, eng.nout), dtype=torch.float64, device=dev)
        out = torch.full((ncol, 4, eng.nout), np.nan, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        mu = 0.8 if call else 1.0
        cols = _cols(d_atm, d_nd, ncol, 82, mu)
        assert L.rhg_synth_lte(eng._h, C.byref(cols), precision, 1,
                               C.cast(C.c_void_p(ref.data_ptr()), C.POINTER(C.c_double))) == 0
        rc = L.rhg_synth_lte_noalloc(eng._h, C.byref(cols), precision,
                                     C.cast(C.c_void_p(out.data_ptr()), C.POINTER(C.c_double)))
        assert rc == 0, L.rhg_last_error()
        torch.cuda.synchronize()
        o, r = out.cpu().numpy(), ref.cpu().numpy()
        print(f"precision {precision} call {call}: {ncol} columns, max |diff| "
              f"{np.max(np.abs(o - r)):.3e}", flush=True)
        np.testing.assert_array_equal(o, r)


def test_noalloc_rejects(eng):
    import torch
    from pyrh_amd.engine import lib
    L = lib()
    dev = torch.device("cuda:0")
    # a batch beyond the scratch: rejected before any launch
    big = 20000
    d_atm = torch.zeros((1, 9, 82), dtype=torch.float64, device=dev)
    d_nd = torch.full((big,), 82, dtype=torch.int32, device=dev)
    out = torch.empty((1,), dtype=torch.float64, device=dev)
    cols = _cols(d_atm, d_nd, big, 82)
    rc = L.rhg_synth_lte_noalloc(eng._h, C.byref(cols), 1,
                                 C.cast(C.c_void_p(out.data_ptr()), C.POINTER(C.c_double)))
    assert rc == RHG_EINVAL and b"scratch" in L.rhg_last_error()
    # a depth count beyond ndep_max: clamped on the device, reported
    atm = W.stokes_map(64, seed=W.SEED_BASE + 95)
    nd = np.full(64, 82, np.int32)
    nd[5] = 90
    d_atm = torch.from_numpy(atm).to(dev)
    d_nd = torch.from_numpy(nd).to(dev)
    out = torch.empty((64, 4, eng.nout), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    cols = _cols(d_atm, d_nd, 64, 82)
    rc = L.rhg_synth_lte_noalloc(eng._h, C.byref(cols), 1,
                                 C.cast(C.c_void_p(out.data_ptr()), C.POINTER(C.c_double)))
    assert rc == RHG_EINVAL and b"ndep" in L.rhg_last_error()
