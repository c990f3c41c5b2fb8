"""Uneven wavelength grids in one launch (rhg_set_grids, rhg_synth_lte_grids).

Every pyrh compute1d call takes its own `wave` (pyrh.pyx:537-541).  A batch
whose columns ask for different grids runs on one context built on the
union of the grids: the columns are sorted by grid and depth count on the
device and each wavefront solves its grid's wavelengths only.  Parity: each
column's spectrum on its grid equals, bit for bit, what a context built on
that grid alone gives for it (the per-wavelength work is the same code on
the same inputs), in the FP64 path and the FP32 form; entries outside a
column's grid are zero.  Batches: the mixed-size run (400 Fe I Zeeman
lines, 2000 wavelengths, 64-256 depths) with overlapping, interleaved,
disjoint and empty grids, and through pyrh_amd.api the 630 nm runs whose
wavelengths go to every solver (DELO-Bezier with and without MO terms,
scalar Bezier, Feautrier)."""
import numpy as np
import pytest

from conftest import GOLDEN
from pyrh_amd import model, workloads as W

pytestmark = pytest.mark.gpu

RUN = GOLDEN / "run_mixed"


@pytest.fixture(scope="module")
def wave():
    return np.load(GOLDEN / "mixed_fixtures.npz")["wave"]


@pytest.fixture(scope="module")
def grids(wave):
    """Five grids on the mixed run's wavelengths: two overlapping spans,
    every seventh point, the red end, and an empty grid."""
    return [wave[0:700], wave[600:1400], wave[::7], wave[1500:], wave[:0]]


@pytest.fixture(scope="module")
def setup(wave, grids):
    from pyrh_amd import api
    from pyrh_amd.engine import Engine
    run = model.open_run(str(RUN))
    merged = np.unique(np.concatenate(grids))
    e = Engine(model.build_model(run, merged))
    pos = api.grid_positions(e.model, grids)
    e.set_grids(pos)
    singles = [Engine(model.build_model(run, g)) if g.size else None for g in grids]
    yield e, pos, singles
    e.close()
    for s in singles:
        if s is not None:
            s.close()


def _check(e, pos, singles, atm, nd, gid, st, mu, prec):
    for g, p in enumerate(pos):
        cols = np.flatnonzero(gid == g)
        outside = np.ones(e.nout, bool)
        outside[p] = False
        assert np.all(st[cols][:, :, outside] == 0.0)
        if cols.size == 0 or singles[g] is None:
            continue
        ref = singles[g].synth_lte(atm[cols], mu=mu, ndep=nd[cols], precision=prec)
        np.testing.assert_array_equal(st[cols][:, :, p], ref)
        print(f"grid {g}: {cols.size} columns x {p.size} wavelengths equal", flush=True)


@pytest.mark.parametrize("prec", (0, 1))
def test_grids_match_single_grid_contexts(setup, prec):
    e, pos, singles = setup
    atm, nd = W.mixed_batch(700, seed=W.SEED_BASE + 40)
    rng = np.random.Generator(np.random.PCG64(W.SEED_BASE + 41))
    gid = rng.integers(0, len(pos), size=atm.shape[0]).astype(np.int32)
    print(f"grids: 700 columns, precision {prec}, sizes {[p.size for p in pos]}", flush=True)
    st = e.synth_lte_grids(atm, gid, mu=0.7, ndep=nd, precision=prec)
    _check(e, pos, singles, atm, nd, gid, st, 0.7, prec)


def test_grids_one_column_and_one_grid(setup):
    """A single column (one block, the rest of the slots padding), and a
    batch all on one grid."""
    e, pos, singles = setup
    atm, nd = W.mixed_batch(130, seed=W.SEED_BASE + 42)
    for a, n, g in ((atm[:1], nd[:1], np.array([2], np.int32)),
                    (atm, nd, np.full(atm.shape[0], 1, np.int32))):
        st = e.synth_lte_grids(a, g, mu=1.0, ndep=n, precision=1)
        _check(e, pos, singles, a, n, g, st, 1.0, 1)


def test_grids_device_pointers(setup):
    """Device buffers (grid ids on the device too): the output starts as
    NaN, and comes back equal to the host-buffer call, zeros included."""
    import torch
    e, pos, singles = setup
    atm, nd = W.mixed_batch(300, seed=W.SEED_BASE + 43)
    gid = (np.arange(atm.shape[0]) % len(pos)).astype(np.int32)[::-1].copy()
    host = e.synth_lte_grids(atm, gid, mu=0.4, ndep=nd, precision=1)
    dev = torch.device("cuda:0")
    d_atm = torch.from_numpy(atm).to(dev)
    d_nd = torch.from_numpy(nd).to(dev)
    d_g = torch.from_numpy(gid).to(dev)
    out = torch.full((atm.shape[0], 4, e.nout), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    e.synth_lte_grids_device(d_atm.data_ptr(), d_nd.data_ptr(), d_g.data_ptr(), atm.shape[0],
                             atm.shape[2], out.data_ptr(), mu=0.4, precision=1)
    np.testing.assert_array_equal(out.cpu().numpy(), host)


def test_grids_out_of_range(setup):
    """A grid id outside [0, ngrid): the host check, and with device buffers
    the device check (the kernels ran on grid 0, inside the buffers); the
    next good call is unaffected."""
    import torch
    from pyrh_amd.engine import EngineError
    e, pos, singles = setup
    atm, nd = W.mixed_batch(200, seed=W.SEED_BASE + 44)
    gid = (np.arange(atm.shape[0]) % 4).astype(np.int32)
    good = e.synth_lte_grids(atm, gid, mu=0.7, ndep=nd)
    dev = torch.device("cuda:0")
    d_atm = torch.from_numpy(atm).to(dev)
    d_nd = torch.from_numpy(nd).to(dev)
    out = torch.empty((atm.shape[0], 4, e.nout), dtype=torch.float64, device=dev)
    for bad in (len(pos), -1):
        g2 = gid.copy()
        g2[77] = bad
        with pytest.raises(EngineError, match="grid out of range"):
            e.synth_lte_grids(atm, g2, mu=0.7, ndep=nd)
        d_g = torch.from_numpy(g2).to(dev)
        torch.cuda.synchronize()
        with pytest.raises(EngineError, match="ndep or grid out of range"):
            e.synth_lte_grids_device(d_atm.data_ptr(), d_nd.data_ptr(), d_g.data_ptr(),
                                     atm.shape[0], atm.shape[2], out.data_ptr(), mu=0.7)
    np.testing.assert_array_equal(e.synth_lte_grids(atm, gid, mu=0.7, ndep=nd), good)


def test_api_synth_grids_matches_synth_batch():
    """pyrh_amd.api.synth_grids on the 630 nm run (three grids: the two Fe I
    lines apart, and a coarse grid over both) against synth_batch per grid."""
    from pyrh_amd import api
    cwd = str(GOLDEN / "run_lte")
    wl = W.wavelengths_630()
    waves = [wl[: wl.size // 2], wl[wl.size // 2:], wl[::5]]
    atm = W.stokes_map(256)
    gid = (np.arange(256) * 7 % 3).astype(np.int32)
    I, Q, U, V, lam, pos = api.synth_grids(cwd, 1.0, 0, atm, waves, gid, precision="fp32")
    for g, w in enumerate(waves):
        cols = np.flatnonzero(gid == g)
        r = api.synth_batch(cwd, 1.0, 0, atm[cols], w, precision="fp32")
        np.testing.assert_array_equal(lam[pos[g]], r[4])
        for a, b in zip((I, Q, U, V), r[:4]):
            np.testing.assert_array_equal(a[cols][:, pos[g]], b)


@pytest.mark.parametrize("run", ("run_lte", "run_lte_field_free", "run_lte_field_free_static",
                                 "run_lte_mo"))
@pytest.mark.parametrize("precision", ("fp64", "fp32"))
def test_api_grids_every_solver(run, precision):
    """Grids whose wavelengths go to every solver: in the FULL_STOKES run,
    one grid mixes the Fe I lines' DELO-Bezier wavelengths with continuum
    points near 600 nm (Feautrier) and one holds only such points; the
    field-free runs solve the scalar Bezier (BEZIER_ANGLE wavelengths); the
    magneto-optical run the DELO-Bezier with MO terms.  Each column on
    its grid against synth_batch on that grid, bit for bit."""
    from pyrh_amd import api
    cwd = str(GOLDEN / run)
    wl = W.wavelengths_630()
    extra = np.linspace(600.0, 601.0, 40)
    if run == "run_lte":
        waves = [np.concatenate([wl[:56], extra[:20]]), wl[56:], extra]
    else:
        waves = [wl[:70], wl[50:], wl[::3]]
    atm = W.stokes_map(192, seed=W.SEED_BASE + 45)
    atm[::2, 3] = 0.0    # half the columns static
    gid = (np.arange(atm.shape[0]) * 5 % len(waves)).astype(np.int32)
    I, Q, U, V, lam, pos = api.synth_grids(cwd, 0.8, 0, atm, waves, gid, precision=precision)
    for g, w in enumerate(waves):
        cols = np.flatnonzero(gid == g)
        r = api.synth_batch(cwd, 0.8, 0, atm[cols], w, precision=precision)
        np.testing.assert_array_equal(lam[pos[g]], r[4])
        for a, b in zip((I, Q, U, V), r[:4]):
            np.testing.assert_array_equal(a[cols][:, pos[g]], b)
    print(f"{run} {precision}: grids {[p.size for p in pos]} equal", flush=True)
