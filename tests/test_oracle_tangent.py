"""Pins of the oracle's tangent-linear build (oracle/dual.h, -DORC_DUAL),
which gives the exact directional derivative (JVP) that the GPU reverse
mode is checked against:

* its values are bit-identical to the double build (same operations, same
  order), for the RHS and for whole differentiated runs;
* its tangents agree with central finite differences of the double build,
  for the RHS on a state with every contravariant velocity bounded away from
  zero (no |U| kink of the Lax-Friedrichs splitting is crossed) and for the
  momentum thickness after a few steps in the amplitudes that move it at
  first order;
* linearity: the tangent of a sum of two directions is the sum of the
  tangents.
"""
import numpy as np
import pytest

from _util import O, cases
from test_oracle_rhs import box_case, exact_grid


def _state(o, rng, n, g, offset=True):
    """A smooth random state on the padded box; with ``offset`` every
    velocity component stays far from zero."""
    ph = rng.uniform(0, 2 * np.pi, size=6)
    Z, Y, X = np.meshgrid(*(np.linspace(0, 2 * np.pi, n + 2 * g),) * 3, indexing="ij")
    rho = 1.1 + 0.2 * np.sin(X + ph[0]) * np.cos(Y + ph[1])
    base = 150.0 if offset else 0.0
    u = base + 30 * np.sin(Y + ph[2])
    v = base + 20 * np.sin(X - Z + ph[3])
    w = base + 15 * np.cos(X + Y + ph[4])
    p = 1e5 * (1 + 0.05 * np.sin(Z + ph[5]))
    return rho, u, v, w, p


def _load(o, prim, dprim=None):
    rho, u, v, w, p = prim
    T = p * cases.MW_AIR / (rho * cases.R_U)
    bv = o.array(0, O.BV)
    vals = (rho, u, v, w, p, T)
    if o.tangent:
        bv[...] = 0
        for l, a in enumerate(vals):
            bv[l, ..., 0] = a
        if dprim is not None:
            drho, du, dv, dw, dp = dprim
            dT = T * (dp / p - drho / rho)
            for l, a in enumerate((drho, du, dv, dw, dp, dT)):
                bv[l, ..., 1] = a
    else:
        for l, a in enumerate(vals):
            bv[l][...] = a
    o.L.orc_cv_from_bv(o.ctx, 0)
    o.L.orc_properties(o.ctx, 0)


def _direction(rng, shape):
    Z, Y, X = np.meshgrid(*(np.linspace(0, 2 * np.pi, shape[0]),) * 3, indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, size=5)
    return (0.01 * np.sin(X + 2 * Y + ph[0]), 3 * np.cos(Y - Z + ph[1]), 2 * np.sin(Z + ph[2]),
            1.5 * np.cos(X + ph[3]), 500 * np.sin(X + Y + Z + ph[4]))


@pytest.mark.parametrize("scheme", [51, 50])
def test_tangent_rhs_values_are_the_double_build(scheme):
    n = 10
    rng = np.random.default_rng(5)
    od, ot = O.Oracle(box_case(n, scheme=scheme)), O.Oracle(box_case(n, scheme=scheme), tangent=True)
    for o in (od, ot):
        exact_grid(o)
    prim = _state(od, rng, n, od.ngg, offset=False)
    _load(od, prim)
    _load(ot, prim, _direction(rng, prim[0].shape))
    od.rhs(0)
    ot.rhs(0)
    np.testing.assert_array_equal(ot.array(0, O.DQ)[..., 0], od.array(0, O.DQ))


@pytest.mark.parametrize("scheme", [51, 50])
@pytest.mark.parametrize("viscous", [0, 2])
def test_tangent_rhs_matches_central_differences(scheme, viscous):
    n = 10
    rng = np.random.default_rng(17 + scheme + viscous)
    case = box_case(n, scheme=scheme, viscous=viscous)
    ot = O.Oracle(case, tangent=True)
    exact_grid(ot)
    prim = _state(ot, rng, n, ot.ngg)
    d = _direction(rng, prim[0].shape)
    _load(ot, prim, d)
    ot.rhs(0)
    jvp = ot.array(0, O.DQ)[..., 1].copy()
    # h = 1e-5 balances round-off (~1e-8 here, falling as 1/h) against the
    # second-order truncation of the energy residual (~1e-8, growing as h^2)
    h = 1e-5

    def dq_at(sign):
        od = O.Oracle(case)
        exact_grid(od)
        _load(od, tuple(a + sign * h * b for a, b in zip(prim, d)))
        od.rhs(0)
        return od.array(0, O.DQ).copy()

    fd = (dq_at(1) - dq_at(-1)) / (2 * h)
    for l in range(5):
        err = np.max(np.abs(fd[l] - jvp[l])) / np.max(np.abs(jvp[l]))
        assert err <= 1e-7, (l, err)


def test_tangent_is_linear_in_the_direction():
    n = 9
    rng = np.random.default_rng(23)
    case = box_case(n)
    out = []
    prim = None
    dirs = []
    for k in range(3):
        ot = O.Oracle(case, tangent=True)
        exact_grid(ot)
        if prim is None:
            prim = _state(ot, rng, n, ot.ngg)
            d1, d2 = _direction(rng, prim[0].shape), _direction(rng, prim[0].shape)
            dirs = [d1, d2, tuple(a + b for a, b in zip(d1, d2))]
        _load(ot, prim, dirs[k])
        ot.rhs(0)
        out.append(ot.array(0, O.DQ)[..., 1].copy())
    scale = np.max(np.abs(out[2]))
    assert np.max(np.abs(out[0] + out[1] - out[2])) <= 1e-12 * scale


def _theta(case, amp, amp_dot=None):
    o = O.Oracle(case, tangent=amp_dot is not None)
    o.setup_mixing_layer()
    o.perturb_fourier(0, case["modes"], amp=amp, amp_dot=amp_dot)
    o.initialize()
    o.step(case["n_steps"])
    return o.momentum_thickness(0, case["theta"])


def test_tangent_run_values_are_the_double_build():
    case = cases.fourier_case(10, n_steps=3)
    a = case["modes"]["amp"]
    th_d, _ = _theta(case, a)
    th_t, _ = _theta(case, a, np.eye(len(a))[0])
    assert th_t == th_d


@pytest.mark.parametrize("mode", [0, 1])
def test_tangent_theta_matches_central_differences(mode):
    """The mean-u and mean-v modes move theta at first order; for them the
    central difference converges (h = 1e-4: truncation and round-off both
    below 1e-7 relative)."""
    case = cases.fourier_case(12, n_steps=4)
    a = case["modes"]["amp"]
    e = np.eye(len(a))[mode]
    _, dth = _theta(case, a, e)
    h = 1e-4
    fd = (_theta(case, a + h * e)[0] - _theta(case, a - h * e)[0]) / (2 * h)
    assert abs(fd - dth) <= 1e-6 * abs(dth), (fd, dth)
