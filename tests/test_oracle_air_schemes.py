"""Pins of the air oracle's characteristic WENO-Z5 and WENO-Z7 (schemes 52, 72) and primitive
reconstruction + Riemann solver path (schemes 1 LF, 3 AUSM+, 4 HLLC) (CPU only).

* L R = I for the air eigenvector pair at a face with equal states;
* F(q, q) = the physical Euler flux for LF, AUSM+ and HLLC;
  supersonic flow takes the upwind flux; HLLC captures a contact exactly;
* free-stream preservation and telescoping conservation of the whole
  residual on a periodic box, for every new scheme and reconstruction.
"""
import numpy as np
import pytest

from _util import O
from test_oracle_rhs import box_case, exact_grid, set_state

GAMMA = 1.4


def _cons(rho, u, p):
    u = np.asarray(u, dtype=float)
    E = p / (GAMMA - 1) + 0.5 * rho * u @ u
    return np.concatenate([[rho], rho * u, [E, p]])


def _flux(rho, u, p, k, jac):
    u = np.asarray(u, dtype=float)
    E = p / (GAMMA - 1) + 0.5 * rho * u @ u
    U = k @ u
    return jac * np.concatenate([[rho * U], rho * u * U + p * k, [(E + p) * U]])


@pytest.mark.parametrize("seed", range(4))
def test_characteristic_matrices_are_inverse(seed):
    rng = np.random.default_rng(seed)
    q = _cons(rng.uniform(0.3, 2.0), rng.uniform(-200, 200, 3), rng.uniform(5e4, 2e5))
    k = rng.normal(size=3)
    L, R = O.char_matrices(q, q, k, k, 0.6, 0.6)
    assert np.max(np.abs(L @ R - np.eye(5))) < 1e-12
    # between different states the Roe average of air is consistent as well
    q2 = _cons(rng.uniform(0.3, 2.0), rng.uniform(-200, 200, 3), rng.uniform(5e4, 2e5))
    L, R = O.char_matrices(q, q2, k, 1.1 * k, 0.6, 0.8)
    assert np.max(np.abs(L @ R - np.eye(5))) < 1e-12


@pytest.mark.parametrize("scheme", [1, 3, 4])
def test_riemann_flux_is_consistent(scheme):
    rng = np.random.default_rng(scheme)
    for _ in range(10):
        rho, u, p = rng.uniform(0.3, 2.0), rng.uniform(-200, 200, 3), rng.uniform(5e4, 2e5)
        k = rng.normal(size=3)
        q = np.concatenate([[rho], u, [p]])
        f = O.riemann(scheme, q, q, k, k, 0.7, 0.7)
        ref = _flux(rho, u, p, k, 0.7)
        assert np.max(np.abs(f - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("scheme", [3, 4])
def test_supersonic_flow_takes_the_upwind_flux(scheme):
    k = np.array([1.0, 0.0, 0.0])
    ql = np.array([1.2, 900.0, 10.0, 0.0, 1e5])
    qr = np.array([0.9, 850.0, -20.0, 5.0, 0.8e5])
    f = O.riemann(scheme, ql, qr, k, k, 1.0, 1.0)
    ref = _flux(ql[0], ql[1:4], ql[4], k, 1.0)
    assert np.max(np.abs(f - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_roe_flux_consistency_upwinding_and_contact():
    """riemannSolver_Roe<Air> (src/RiemannSolver.cuh:279-417): F(q, q) is the
    physical flux for any entropy-fix delta; with every wave speed above the
    delta the Roe property A_roe dQ = dF makes it the upwind flux; with no
    entropy fix a stationary contact keeps zero mass flux and the pressure."""
    rng = np.random.default_rng(11)
    for _ in range(10):
        rho, u, p = rng.uniform(0.3, 2.0), rng.uniform(-200, 200, 3), rng.uniform(5e4, 2e5)
        k = rng.normal(size=3)
        q = np.concatenate([[rho], u, [p]])
        f = O.roe(q, q, k, k, 0.7, 0.7, 50.0, 70.0)
        ref = _flux(rho, u, p, k, 0.7)
        assert np.max(np.abs(f - ref)) <= 1e-12 * np.max(np.abs(ref))
    k = np.array([1.0, 0.0, 0.0])
    ql = np.array([1.2, 900.0, 10.0, 0.0, 1e5])
    qr = np.array([0.9, 850.0, -20.0, 5.0, 0.8e5])
    f = O.roe(ql, qr, k, k, 1.0, 1.0, 1.0, 1.0)
    ref = _flux(ql[0], ql[1:4], ql[4], k, 1.0)
    assert np.max(np.abs(f - ref)) <= 1e-11 * np.max(np.abs(ref))
    k = np.array([0.0, 0.6, 0.8])
    ql = np.array([1.3, 0.0, 0.0, 0.0, 1e5])
    qr = np.array([0.4, 0.0, 0.0, 0.0, 1e5])
    f = O.roe(ql, qr, k, k, 1.0, 1.0)
    np.testing.assert_allclose(f, [0.0, 0.0, 0.6e5, 0.8e5, 0.0], atol=1e-9 * 1e5)
    # with the entropy fix the contact wave (speed 0) is smoothed: mass diffuses
    assert abs(O.roe(ql, qr, k, k, 1.0, 1.0, 10.0, 10.0)[0]) > 0


def test_hllc_resolves_an_isolated_contact():
    k = np.array([0.0, 0.6, 0.8])
    for un in (35.0, -35.0):
        u = un * k
        ql = np.concatenate([[1.3], u, [1e5]])
        qr = np.concatenate([[0.4], u, [1e5]])
        f = O.riemann(4, ql, qr, k, k, 1.0, 1.0)
        q = ql if un > 0 else qr
        ref = _flux(q[0], q[1:4], q[4], k, 1.0)
        assert np.max(np.abs(f - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_lax_friedrichs_dissipation_is_the_largest_spectral_radius():
    """LF: F = (F(ql) + F(qr))/2 - lambda/2 J (Q(qr) - Q(ql)), lambda the
    larger |U| + c |grad k| of the two states."""
    k = np.array([0.3, -0.4, 1.2])
    ql = np.array([1.1, 50.0, 20.0, -10.0, 1.1e5])
    qr = np.array([0.8, -30.0, 5.0, 40.0, 0.9e5])
    f = O.riemann(1, ql, qr, k, k, 0.9, 0.9)
    g = np.linalg.norm(k)
    lam = max(abs(k @ q[1:4]) + np.sqrt(GAMMA * q[4] / q[0]) * g for q in (ql, qr))
    Ql, Qr = _cons(ql[0], ql[1:4], ql[4])[:5], _cons(qr[0], qr[1:4], qr[4])[:5]
    ref = 0.5 * (_flux(ql[0], ql[1:4], ql[4], k, 0.9) + _flux(qr[0], qr[1:4], qr[4], k, 0.9)) \
        - 0.5 * 0.9 * lam * (Qr - Ql)
    assert np.max(np.abs(f - ref)) <= 1e-12 * np.max(np.abs(ref))


def _case(scheme, recon=2, n=10):
    c = box_case(n, scheme=scheme)
    c["reconstruction"] = recon
    return c


@pytest.mark.parametrize("scheme,recon", [(52, 2), (72, 2), (1, 1), (1, 2), (2, 2), (3, 2), (3, 3), (4, 1), (4, 2),
                                          (4, 3)])
def test_free_stream_and_conservation(scheme, recon):
    o = O.Oracle(_case(scheme, recon))
    exact_grid(o)
    set_state(o, 1.2, 80.0, -30.0, 15.0, 9.0e4)
    o.rhs(0)
    dq = o.array(0, O.DQ)
    cv = o.array(0, O.CV)
    for l in range(5):
        assert np.max(np.abs(dq[l])) <= 1e-13 * np.max(np.abs(cv[l]))
    # periodic state of period n-1 nodes, dq summed over distinct nodes
    n, g = 10, o.ngg
    idx = np.arange(-g, n + g)
    ph = 2 * np.pi * (idx % (n - 1)) / (n - 1)
    Z, Y, X = np.meshgrid(ph, ph, ph, indexing="ij")
    rho = 1.2 + 0.1 * np.sin(X + 2 * Y) * np.cos(Z)
    set_state(o, rho, 50 + 20 * np.cos(Y) + 5 * np.sin(X + Z), 10 * np.sin(Z - X), -8 * np.cos(2 * Y + Z),
              1e5 + 3e3 * np.sin(X) * np.sin(Y))
    L = O.lib()
    L.orc_apply_bc(o.ctx, 0)
    L.orc_properties(o.ctx, 0)
    o.rhs(0)
    dq = o.array(0, O.DQ)[:, :-1, :-1, :-1]
    tot = dq.sum(axis=(1, 2, 3))
    scale = np.abs(dq).sum(axis=(1, 2, 3))
    assert np.all(np.abs(tot) <= 1e-13 * scale), (tot, scale)
    assert np.all(scale > 0)




def _wave_dq(scheme, n):
    """dq of a small-amplitude plane wave along (1, 2, -1) in index space,
    n - 1 nodes per period on every axis."""
    o = O.Oracle(_case(scheme, 2, n))
    exact_grid(o)
    g = o.ngg
    idx = np.arange(-g, n + g)
    ph = 2 * np.pi * (idx % (n - 1)) / (n - 1)
    Z, Y, X = np.meshgrid(ph, ph, ph, indexing="ij")
    W = X + 2 * Y - Z
    set_state(o, 1.2 * (1 + 1e-2 * np.sin(W)), 40.0 + 0.5 * np.cos(W), -20.0 + 0.3 * np.sin(W + 1),
              10.0 + 0 * W, 1e5 * (1 + 1e-2 * np.sin(W + 0.3)))
    L = O.lib()
    L.orc_apply_bc(o.ctx, 0)
    L.orc_properties(o.ctx, 0)
    o.rhs(0)
    return o.array(0, O.DQ).copy()


def test_characteristic_and_componentwise_weno_converge_together():
    """In the smooth limit WENO-Z weights tend to the linear ones, where a
    per-face linear projection commutes with the reconstruction: the
    characteristic residuals (52, 72) approach the component-wise ones (51,
    71) at the schemes' order.  A stencil shifted by one point would sit at
    O(h k) ~ 0.5 instead."""
    err = {}
    for n in (20, 30):
        d = {s: _wave_dq(s, n) for s in (51, 52, 71, 72)}
        for l in range(5):
            scale = np.max(np.abs(d[71][l]))
            e72 = np.max(np.abs(d[72][l] - d[71][l])) / scale
            e52 = np.max(np.abs(d[52][l] - d[51][l])) / scale
            err[(n, l)] = (e72, e52)
            assert e72 <= 5e-3 and e52 <= 2e-2, (n, l, e72, e52)
            assert np.max(np.abs(np.roll(d[72][l], 1, axis=1) - d[71][l])) / scale > 0.2
    for l in range(5):
        assert err[(30, l)][0] <= 0.1 * err[(20, l)][0], (l, err)
        assert err[(30, l)][1] <= 0.3 * err[(20, l)][1], (l, err)
