"""Pins of the oracle's sixth-order energy-preserving scheme (inviscid_scheme
6, compute_convective_term_ep, src/InviscidScheme.cu:550-768; Pirozzoli,
J. Comput. Phys. 230 (2011) 2997-3014), air:

* a uniform free stream gives dq = 0;
* on an xyz-periodic box the residual telescopes: the sum of dq over the
  distinct nodes is 0 for every conserved variable;
* on a smooth periodic wave along one axis (Cartesian grid, J = h^3) the
  residual converges to -J dF/dx of the exact fluxes at sixth order (rate
  above 5.5 from 16 to 32 nodes per period);
* the split form conserves kinetic energy: on a periodic box at uniform
  pressure sum_i (u . dq_m - |u|^2/2 dq_rho) = 0 to round-off, the property
  the scheme is named for.
"""
import numpy as np
import pytest

from _util import O, cases
from test_oracle_rhs import box_case, exact_grid, set_state

GAMMA = 1.4


def _periodic_box(n, scheme=6):
    o = O.Oracle(box_case(n, scheme=scheme, viscous=0))
    return o


def test_free_stream_gives_zero():
    o = _periodic_box(9)
    exact_grid(o)
    set_state(o, 1.2, 80.0, -30.0, 15.0, 9.0e4)
    o.rhs(0)
    assert np.all(o.array(0, O.DQ) == 0.0)


def _state_on_box(o, n, fields):
    g = o.ngg
    idx = np.arange(-g, n + g)
    ph = 2 * np.pi * (idx % (n - 1)) / (n - 1)
    Z, Y, X = np.meshgrid(ph, ph, ph, indexing="ij")
    set_state(o, *fields(X, Y, Z))
    L = O.lib()
    L.orc_apply_bc(o.ctx, 0)
    L.orc_properties(o.ctx, 0)


def test_periodic_box_telescopes():
    n = 12
    o = _periodic_box(n)
    exact_grid(o)
    _state_on_box(o, n, lambda X, Y, Z: (1.2 + 0.1 * np.sin(X + 2 * Y) * np.cos(Z), 50 + 20 * np.cos(Y),
                                         10 * np.sin(Z - X), -8 * np.cos(2 * Y + Z),
                                         1e5 + 3e3 * np.sin(X) * np.sin(Y)))
    o.rhs(0)
    dq = o.array(0, O.DQ)[:, :-1, :-1, :-1]
    tot = dq.sum(axis=(1, 2, 3))
    scale = np.abs(dq).sum(axis=(1, 2, 3))
    assert np.all(np.abs(tot) <= 1e-13 * scale), (tot, scale)


def _wave_error(n, h=0.01):
    """max |dq + J dF/dx| over the nodes of a smooth x-periodic wave."""
    o = _periodic_box(n)
    exact_grid(o, h)
    g = o.ngg
    L = (n - 1) * h
    i = np.arange(-g, n + g)
    x = (i % (n - 1)) * h
    k = 2 * np.pi / L
    rho1 = 1.0 + 0.2 * np.sin(k * x)
    u1 = 60.0 + 25.0 * np.cos(k * x)
    p1 = 9e4 + 4e3 * np.sin(2 * k * x)
    shape = (n + 2 * g,) * 3
    bc = lambda a: np.broadcast_to(a[None, None, :], shape).copy()
    set_state(o, bc(rho1), bc(u1), np.zeros(shape), np.zeros(shape), bc(p1))
    Lb = O.lib()
    Lb.orc_apply_bc(o.ctx, 0)
    o.rhs(0)
    dq = o.array(0, O.DQ)[:, 0, 0, :]
    xs = np.arange(n) * h
    r = 1.0 + 0.2 * np.sin(k * xs)
    u = 60.0 + 25.0 * np.cos(k * xs)
    p = 9e4 + 4e3 * np.sin(2 * k * xs)
    dr, du, dp = 0.2 * k * np.cos(k * xs), -25.0 * k * np.sin(k * xs), 8e3 * k * np.cos(2 * k * xs)
    E = p / (GAMMA - 1) + 0.5 * r * u * u
    dE = dp / (GAMMA - 1) + 0.5 * dr * u * u + r * u * du
    # exact x-derivatives of the flux F = (rho u, rho u^2 + p, 0, 0, (E + p) u): with the metric 1/h and
    # J = h^3 the face fluxes are h^2 F, so dq = -(h^2 F)' h = -J dF/dx
    dF = np.array([dr * u + r * du, dr * u * u + 2 * r * u * du + dp, 0 * xs, 0 * xs, (dE + dp) * u + (E + p) * du])
    J = h ** 3
    err = np.max(np.abs(dq + J * dF), axis=1)
    scale = np.max(np.abs(J * dF), axis=1)
    return err[[0, 1, 4]] / scale[[0, 1, 4]]


def test_sixth_order_on_a_smooth_wave():
    e1 = _wave_error(17)
    e2 = _wave_error(33)
    rate = np.log2(e1 / e2)
    print("errors", e1, e2, "rates", rate)
    assert np.all(rate > 5.5), rate


def test_kinetic_energy_is_conserved_by_the_split_form():
    n = 12
    o = _periodic_box(n)
    exact_grid(o)
    _state_on_box(o, n, lambda X, Y, Z: (1.2 + 0.2 * np.sin(X + 2 * Y) * np.cos(Z), 50 * np.sin(Y) * np.cos(Z),
                                         30 * np.sin(Z - X), -20 * np.cos(2 * Y + X),
                                         np.full(X.shape, 1e5)))
    o.rhs(0)
    g = o.ngg
    dq = o.array(0, O.DQ)[:, :-1, :-1, :-1]
    bv = o.array(0, O.BV)[:, g:-g - 1, g:-g - 1, g:-g - 1]
    u, v, w = bv[1], bv[2], bv[3]
    ke = u * dq[1] + v * dq[2] + w * dq[3] - 0.5 * (u * u + v * v + w * w) * dq[0]
    terms = np.abs(u * dq[1]).sum() + np.abs(0.5 * (u * u + v * v + w * w) * dq[0]).sum()
    print("kinetic energy rate", ke.sum(), "scale", terms)
    assert abs(ke.sum()) <= 1e-12 * terms
