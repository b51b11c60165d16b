"""Reverse mode on the GPU against the oracle's tangent-linear build (needs a GPU).

* RHS VJP, dot-product test: for a random dq cotangent w and a state
  direction d, <w, J d> from the oracle's dual numbers must equal
  <J^T w, d> from ostc_rhs_vjp, to 1e-11 relative (both are exact
  derivatives; they differ by round-off only);
* the reverse sweep of a differentiated run (Fourier-mode amplitudes ->
  momentum thickness after n SSP-RK3 steps) against the tangent oracle run
  once per amplitude, component by component;
* the same gradient against central finite differences of the GPU forward
  run with the step 1e-6, all 32 amplitudes, norm-wise within 1e-6 of the
  largest gradient component;
* checkpoint spacing changes the recomputation, never the result (bitwise).
"""
import numpy as np
import pytest

from _util import O, cases, pkg
from test_oracle_rhs import box_case, exact_grid
from test_oracle_tangent import _direction, _load, _state, _theta

pytestmark = pytest.mark.gpu


def _gpu_from_oracle(case, o):
    s = pkg.Solver(case)
    s.set_geometry(0, o.array(0, O.JAC)[0], o.array(0, O.METRIC))
    s.set(0, pkg.FIELD_CV, o.array(0, O.CV)[..., 0])
    s.set(0, pkg.FIELD_BV, o.array(0, O.BV)[..., 0])
    s.set(0, pkg.FIELD_MUL, o.array(0, O.MUL)[..., 0])
    return s


@pytest.mark.parametrize("scheme", [51, 50])
@pytest.mark.parametrize("viscous", [0, 2])
@pytest.mark.parametrize("offset", [True, False], ids=["moving", "at-rest-mean"])
def test_rhs_vjp_dot_product(scheme, viscous, offset):
    n = 12
    print(f"[gpu] rhs vjp dot product scheme {scheme} viscous {viscous} offset {offset}", flush=True)
    rng = np.random.default_rng(100 + scheme + viscous + offset)
    case = box_case(n, scheme=scheme, viscous=viscous)
    ot = O.Oracle(case, tangent=True)
    exact_grid(ot)
    prim = _state(ot, rng, n, ot.ngg, offset=offset)
    _load(ot, prim, _direction(rng, prim[0].shape))
    ot.rhs(0)
    jvp = ot.array(0, O.DQ)[..., 1]
    dcv = ot.array(0, O.CV)[..., 1]
    s = _gpu_from_oracle(case, ot)
    for trial in range(3):
        w = rng.normal(size=jvp.shape)
        cv_bar = s.rhs_vjp(0, w)
        lhs = np.sum(w * jvp)
        rhs = np.sum(cv_bar * dcv)
        scale = np.sum(np.abs(w * jvp))
        print(f"  trial {trial}: <w,Jd> {lhs:.16e} <J^T w,d> {rhs:.16e} rel {abs(lhs - rhs) / scale:.2e}", flush=True)
        assert abs(lhs - rhs) <= 1e-11 * scale
    s.close()


def test_rhs_vjp_of_one_cell_cotangent_by_components():
    """w = 1 on one variable of one cell: <J^T w, d> is that cell's JVP
    entry, for cells on the boundary, next to it and inside."""
    n = 10
    rng = np.random.default_rng(3)
    case = box_case(n)
    ot = O.Oracle(case, tangent=True)
    exact_grid(ot)
    prim = _state(ot, rng, n, ot.ngg, offset=False)
    _load(ot, prim, _direction(rng, prim[0].shape))
    ot.rhs(0)
    jvp = ot.array(0, O.DQ)[..., 1]
    dcv = ot.array(0, O.CV)[..., 1]
    s = _gpu_from_oracle(case, ot)
    for (l, k, j, i) in [(0, 0, 0, 0), (4, 0, 5, 9), (1, 1, 1, 2), (2, 5, 4, 3), (3, 9, 9, 9)]:
        w = np.zeros(jvp.shape)
        w[l, k, j, i] = 1.0
        got = np.sum(s.rhs_vjp(0, w) * dcv)
        assert abs(got - jvp[l, k, j, i]) <= 1e-11 * np.max(np.abs(jvp[l])), (l, k, j, i, got, jvp[l, k, j, i])
    s.close()


_GRAD = {}


def _gpu_gradient(case, every):
    s = pkg.Solver(case)
    s.setup_fourier_case()
    th = s.adjoint_forward(case["n_steps"], every, case["theta"])
    g = s.adjoint_backward(case["modes"])
    s.close()
    return th, g


def _tangent_gradient(case):
    key = (case["blocks"][0]["mx"], case["n_steps"])
    if key not in _GRAD:
        a = case["modes"]["amp"]
        th, g = None, np.zeros(len(a))
        for m in range(len(a)):
            th, g[m] = _theta(case, a, np.eye(len(a))[m])
        _GRAD[key] = (th, g)
    return _GRAD[key]


def test_gradient_matches_the_tangent_oracle():
    case = cases.fourier_case(12, n_steps=7, checkpoint_every=3)
    print("[gpu] adjoint gradient, 12^3, 7 steps, checkpoints every 3", flush=True)
    th_g, g = _gpu_gradient(case, 3)
    th_o, g_o = _tangent_gradient(case)
    print(f"  theta gpu {th_g:.16e} oracle {th_o:.16e}", flush=True)
    err = np.abs(g - g_o)
    print(f"  max |g - g_tangent| / max|g| = {np.max(err) / np.max(np.abs(g_o)):.2e}", flush=True)
    assert abs(th_g - th_o) <= 1e-12 * abs(th_o)
    assert np.all(err <= 1e-9 * np.max(np.abs(g_o)) + 1e-7 * np.abs(g_o)), (g, g_o)


def test_checkpoint_spacing_does_not_change_the_gradient():
    case = cases.fourier_case(12, n_steps=5)
    print("[gpu] adjoint gradient with checkpoints every 1, 2, 5 steps", flush=True)
    ref = _gpu_gradient(case, 5)
    for every in (1, 2):
        th, g = _gpu_gradient(case, every)
        assert th == ref[0]
        np.testing.assert_array_equal(g, ref[1])


@pytest.mark.parametrize("n,steps", [(12, 4), (24, 6)])
def test_gradient_matches_central_differences(n, steps):
    """All 32 amplitudes against FP64 central differences of the GPU forward
    run at h = 1e-6: max_m |fd_m - g_m| <= 1e-6 max_m |g_m|."""
    case = cases.fourier_case(n, n_steps=steps, checkpoint_every=2)
    print(f"[gpu] adjoint vs central differences, all amplitudes, {n}^3 x {steps} steps", flush=True)
    _, g = _gpu_gradient(case, 2)
    a = np.array(case["modes"]["amp"])
    h = 1e-6
    s = pkg.Solver(case)

    def theta(amp):
        s.setup_mixing_layer()
        s.perturb_fourier(0, case["modes"], amp)
        s.initialize()
        s.rk3(steps, pkg.MODE_BATCHED, want_dt=False)
        return s.momentum_thickness(0, case["theta"])

    fd = np.array([(theta(a + h * e) - theta(a - h * e)) / (2 * h) for e in np.eye(len(a))])
    s.close()
    worst = np.max(np.abs(fd - g)) / np.max(np.abs(g))
    print(f"  |fd - g|_inf / |g|_inf = {worst:.3e} (mode {int(np.argmax(np.abs(fd - g)))})", flush=True)
    assert worst <= 1e-6


def test_reverse_sweep_rejects_what_it_does_not_cover():
    print("[gpu] adjoint error paths", flush=True)
    multi = cases.small_multiblock_case()
    multi["fixed_time_step"], multi["dt"] = 1, 1e-9
    s = pkg.Solver(multi)
    s.setup_mixing_layer()
    s.initialize()
    t = cases.fourier_case(12)["theta"]
    with pytest.raises(pkg.OstcError, match="single block"):
        s.adjoint_forward(2, 1, t)
    s.close()
    case = cases.single_block_case(12)
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    s.initialize()
    with pytest.raises(pkg.OstcError, match="fixed_time_step"):
        s.adjoint_forward(2, 1, t)
    with pytest.raises(pkg.OstcError, match="has not run"):
        s.adjoint_backward()
    s.close()


def test_torch_autograd_wrapper_returns_the_adjoint_gradient():
    """autodiff.momentum_thickness: theta and amp.grad from torch's backward()
    are the library's forward and reverse sweeps, bit for bit."""
    import torch

    device = "cpu"
    case = cases.fourier_case(12, n_steps=4, checkpoint_every=2)
    print("[gpu] torch autograd wrapper", flush=True)
    th_ref, g_ref = _gpu_gradient(case, 2)
    s = pkg.Solver(case)
    amp = torch.tensor(case["modes"]["amp"], dtype=torch.float64, device=device, requires_grad=True)
    th = pkg.autodiff.momentum_thickness(amp, s)
    (2.0 * th).backward()
    assert th.item() == th_ref
    np.testing.assert_array_equal(amp.grad.cpu().numpy(), 2.0 * g_ref)
    s.close()
