"""pyrh_amd.autograd.stokes: the synthesis as a torch.autograd.Function whose
backward is the library's adjoint.  Checked against the response functions
of the same engine, and against central differences of a scalar loss of the
plain FP64 synthesis."""
import numpy as np
import pytest

from pyrh_amd import model, workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng(setups):
    from pyrh_amd.engine import Engine
    e = Engine(model.build_model(setups["lte"], W.wavelengths_630()))
    yield e
    e.close()


@pytest.mark.parametrize("precision", (0, 1))
def test_autograd_backward_is_the_adjoint(eng, precision):
    import torch
    from pyrh_amd import autograd
    from pyrh_amd.engine import RF_PARAMS, RF_ROWS
    atm = W.stokes_map(24, seed=W.SEED_BASE + 81)
    x = torch.from_numpy(atm).cuda().requires_grad_(True)
    st = autograd.stokes(eng, x, mu=1.0, precision=precision)
    ref = eng.synth_lte(atm, precision=precision)
    # the forward runs the adjoint's sweep, which synthesises the same Stokes
    np.testing.assert_allclose(st.detach().cpu().numpy(), ref, rtol=1e-12, atol=0)
    w = torch.linspace(-1.0, 1.0, st.numel(), dtype=torch.float64, device="cuda").view_as(st)
    (st * w).sum().backward()
    g = x.grad.cpu().numpy()
    _, rf = eng.rf_lte(atm, precision=precision)           # [ncol, 6, nd, 4, nout]
    wn = w.cpu().numpy()
    expect = np.einsum("csl,cpksl->cpk", wn, rf.astype(np.float64))
    for ip, row in enumerate(RF_ROWS):
        scale = np.max(np.abs(expect[:, ip]))
        err = np.max(np.abs(g[:, row] - expect[:, ip])) / scale
        print(f"precision {precision} {RF_PARAMS[ip]}: max err {err:.2e}", flush=True)
        assert err < 1e-10
    for row in (0, 2, 8):
        assert np.all(g[:, row] == 0.0)


def test_autograd_against_finite_differences(eng):
    """d L / d T(k) for L = sum of I and d L / d B(k) for L = sum of |V|
    (weights sign(V), held fixed) against central differences of the FP64
    synthesis at depths of the line-forming region, relative to the largest
    gradient of the row.  Weights without cancellation keep the finite
    difference meaningful: the synthesis has kinks (Fritsch-Butland limiter,
    control-point clamps) and Humlicek region boundaries."""
    import torch
    from pyrh_amd import autograd
    atm = W.stokes_map(4, seed=W.SEED_BASE + 82)
    base = eng.synth_lte(atm)
    for row, h, wn in ((1, 0.01, np.zeros_like(base)), (5, 0.1, np.zeros_like(base))):
        if row == 1:
            wn[:, 0] = 1.0
        else:
            wn[:, 3] = np.sign(base[:, 3])
        x = torch.from_numpy(atm).cuda().requires_grad_(True)
        st = autograd.stokes(eng, x, precision=0)
        (st * torch.from_numpy(wn).cuda()).sum().backward()
        g = x.grad.cpu().numpy()
        loss = lambda a: float(np.sum(wn * eng.synth_lte(a)))
        gmax = float(np.max(np.abs(g[:, row].sum(axis=0))))
        for k in (45, 55, 62):
            p, m = atm.copy(), atm.copy()
            p[:, row, k] += h
            m[:, row, k] -= h
            fd = (loss(p) - loss(m)) / (2 * h)
            ad = float(np.sum(g[:, row, k]))
            print(f"row {row} depth {k}: adjoint {ad:.6e}  fd {fd:.6e}  "
                  f"diff/max {abs(ad - fd) / gmax:.2e}", flush=True)
            assert abs(ad - fd) <= 1e-3 * gmax
