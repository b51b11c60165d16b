"""Pins of the oracle's whole RHS, RK3 update, time step and properties:

* free-stream preservation: a uniform state on an exactly representable grid
  gives dq == 0 exactly;
* telescoping conservation: on a box periodic in x, y and z the residual
  summed over the distinct nodes vanishes to round-off;
* a viscous closed form: u = a*y^2 gives (momentum-x) dq = 2*a*mu*dx*dy*dz;
* SSP-RK3 with dq = lambda*J*cv amplifies cv by 1 + z + z^2/2 + z^3/6;
* dt of a uniform state at rest and Sutherland's law at its reference point;
* an independent numpy restatement of the RHS (tests/numpy_rhs.py) on tiny
  grids with random smooth and rough states;
* a golden fixture of the oracle's dq (tests/golden/), guarding the oracle
  against silent drift.
"""
import os

import numpy as np
import pytest

import numpy_rhs
from _util import O, ROOT, cases

GOLDEN = os.path.join(ROOT, "tests", "golden", "oracle_dq_single12.npz")


def box_case(n=10, faces=None, scheme=51, viscous=2):
    """A single block of n^3 nodes; faces: dict axis -> boundary type."""
    faces = faces or {0: cases.BC_PERIODIC, 1: cases.BC_PERIODIC, 2: cases.BC_PERIODIC}
    st = cases.air_streams()
    return dict(n_var=5, inviscid_scheme=scheme, viscous_order=viscous, cfl=0.5, prandtl=0.72,
                fixed_time_step=0, dt=0.0, rho_ref=st["upper"][0], v_ref=abs(st["upper"][1]), streams=st,
                blocks=[dict(mx=n, my=n, mz=n, gi0=0, gj0=0, gk0=0, x0=0.0, y0=0.0, z0=0.0)],
                boundaries=[dict(block=0, axis=a, side=s, type=t) for a, t in faces.items() for s in (-1, 1)],
                inner_faces=[], n_cells=n ** 3)


def exact_grid(o, h=0.5):
    """Uniform grid with exactly representable coordinates: metric = 1/h, J = h^3."""
    L = O.lib()
    L.orc_uniform_grid(o.ctx, 0, 0.0, 0.0, 0.0, h, h, h)
    assert L.orc_compute_metric(o.ctx, 0) == 0


def set_state(o, rho, u, v, w, p):
    bv = o.array(0, O.BV)
    T = p * cases.MW_AIR / (rho * cases.R_U)
    for l, a in enumerate((rho, u, v, w, p, T)):
        bv[l][...] = a
    L = O.lib()
    L.orc_cv_from_bv(o.ctx, 0)
    L.orc_properties(o.ctx, 0)


def test_free_stream_is_preserved_exactly():
    o = O.Oracle(box_case(9))
    exact_grid(o)
    set_state(o, 1.2, 80.0, -30.0, 15.0, 9.0e4)
    o.rhs(0)
    assert np.all(o.array(0, O.DQ) == 0.0)


@pytest.mark.parametrize("scheme", [51, 50, 71])
def test_periodic_box_conserves(scheme):
    n = 12
    o = O.Oracle(box_case(n, scheme=scheme))
    exact_grid(o)
    g = o.ngg
    rng = np.random.default_rng(7)
    # a periodic state of period n-1 nodes (node n-1 repeats node 0)
    idx = np.arange(-g, n + g)
    ph = 2 * np.pi * (idx % (n - 1)) / (n - 1)
    Z, Y, X = np.meshgrid(ph, ph, ph, indexing="ij")
    rho = 1.2 + 0.1 * np.sin(X + 2 * Y) * np.cos(Z)
    u = 50 + 20 * np.cos(Y) + 5 * np.sin(X + Z)
    v = 10 * np.sin(Z - X)
    w = -8 * np.cos(2 * Y + Z)
    p = 1e5 + 3e3 * np.sin(X) * np.sin(Y)
    set_state(o, rho, u, v, w, p)
    L = O.lib()
    L.orc_apply_bc(o.ctx, 0)
    L.orc_properties(o.ctx, 0)
    o.rhs(0)
    dq = o.array(0, O.DQ)[:, :-1, :-1, :-1]      # distinct nodes only
    tot = dq.sum(axis=(1, 2, 3))
    scale = np.abs(dq).sum(axis=(1, 2, 3))
    assert np.all(np.abs(tot) <= 1e-13 * scale), (tot, scale)
    del rng


def test_viscous_quadratic_profile_closed_form():
    n = 10
    h = 0.5
    o = O.Oracle(box_case(n, faces={0: cases.BC_PERIODIC, 1: cases.BC_OUTFLOW, 2: cases.BC_PERIODIC}))
    exact_grid(o, h)
    g = o.ngg
    j = np.arange(-g, n + g) * h
    a = 3.0
    Y = np.broadcast_to(j[None, :, None], (n + 2 * g,) * 3)
    set_state(o, np.full(Y.shape, 1.1), a * Y * Y, np.zeros(Y.shape), np.zeros(Y.shape), np.full(Y.shape, 9e4))
    mu = o.array(0, O.MUL)[0, 0, 0, 0]
    L = O.lib()
    L.orc_zero_dq(o.ctx, 0)
    L.orc_viscous(o.ctx, 0)
    dq = o.array(0, O.DQ)
    expect = 2 * a * mu * h ** 3
    np.testing.assert_allclose(dq[1], expect, rtol=1e-12)
    assert np.all(dq[0] == 0)


def test_rk3_amplification_polynomial():
    n = 8
    o = O.Oracle(box_case(n))
    exact_grid(o)
    rng = np.random.default_rng(3)
    cv = o.array(0, O.CV)
    cv[...] = rng.uniform(0.5, 2.0, size=cv.shape)
    qn0 = cv.copy()
    jac = o.array(0, O.JAC)[0]
    g = o.ngg
    lam, dt = -0.7, 1.3
    L = O.lib()
    L.orc_copy_qn(o.ctx, 0)
    for rk in range(3):
        dq = o.array(0, O.DQ)
        dq[...] = lam * (jac * cv)[:, g:-g, g:-g, g:-g]
        L.orc_update(o.ctx, 0, dt, rk)
    z = lam * dt
    amp = 1 + z + z * z / 2 + z ** 3 / 6
    np.testing.assert_allclose(cv[:, g:-g, g:-g, g:-g], amp * qn0[:, g:-g, g:-g, g:-g], rtol=1e-14)


def test_local_dt_of_uniform_state_at_rest():
    h = 0.5
    o = O.Oracle(box_case(8))
    exact_grid(o, h)
    rho, p = 1.3, 8e4
    set_state(o, rho, 0.0, 0.0, 0.0, p)
    o.local_dt(0)
    T = p * cases.MW_AIR / (rho * cases.R_U)
    c = np.sqrt(1.4 * cases.R_U * T / cases.MW_AIR)
    mu = cases.sutherland(T)
    expect = 0.5 / (3 * c / h + 3 / h ** 2 * 1.4 / rho * mu / 0.72)
    np.testing.assert_allclose(o.array(0, O.DT_LOCAL), expect, rtol=1e-14)
    assert o.global_dt() == pytest.approx(expect, rel=1e-14)


def test_sutherland_reference_point_and_law():
    assert O.lib().orc_sutherland(288.16) == 1.7894e-5
    t1, t2 = 250.0, 900.0
    r = O.lib().orc_sutherland(t2) / O.lib().orc_sutherland(t1)
    assert r == pytest.approx((t2 / t1) ** 1.5 * (t1 + 110) / (t2 + 110), rel=1e-14)


@pytest.mark.parametrize("rough", [False, True])
def test_rhs_matches_independent_numpy_restatement(rough):
    n = 11
    o = O.Oracle(box_case(n, faces={0: cases.BC_PERIODIC, 1: cases.BC_OUTFLOW, 2: cases.BC_PERIODIC}))
    L = O.lib()
    L.orc_uniform_grid(o.ctx, 0, 0.1, -0.3, 0.2, 0.37e-4, 0.52e-4, 0.29e-4)
    assert L.orc_compute_metric(o.ctx, 0) == 0
    g = o.ngg
    rng = np.random.default_rng(11 + rough)
    shp = (n + 2 * g,) * 3
    ph = rng.uniform(0, 2 * np.pi, size=6)
    Z, Y, X = np.meshgrid(*(np.linspace(0, 2 * np.pi, n + 2 * g),) * 3, indexing="ij")
    rho = 1.1 + 0.2 * np.sin(X + ph[0]) * np.cos(Y + ph[1])
    u = 100 * np.tanh(3 * np.sin(Y + ph[2])) + 10 * np.cos(Z)
    v = 20 * np.sin(X - Z + ph[3])
    w = 15 * np.cos(X + Y + ph[4])
    p = 1e5 * (1 + 0.05 * np.sin(Z + ph[5]))
    if rough:
        rho = rho * (1 + 0.3 * (rng.uniform(size=shp) > 0.8))
        u = u + 40 * rng.normal(size=shp)
    set_state(o, rho, u, v, w, p)
    o.rhs(0)
    c = o.case
    ref = numpy_rhs.rhs(o.array(0, O.CV), o.array(0, O.BV), o.array(0, O.MUL)[0], o.array(0, O.JAC)[0],
                        o.array(0, O.METRIC), g, (c["rho_ref"] * c["v_ref"]) ** 2, c["v_ref"], c["prandtl"])
    got = o.array(0, O.DQ)
    for l in range(5):
        err = np.max(np.abs(got[l] - ref[l])) / np.max(np.abs(ref[l]))
        assert err <= 1e-11, (l, err)


def _golden_case():
    return cases.single_block_case(12)


def make_golden():
    """Writes the golden fixture (run by tests/golden/make_golden.py)."""
    case = _golden_case()
    o = O.Oracle(case)
    o.setup_mixing_layer()
    o.initialize()
    o.rhs(0)
    dq0 = o.array(0, O.DQ).copy()
    dt = o.step(3)
    np.savez_compressed(GOLDEN, dq=dq0, bv_after3=o.array(0, O.BV).copy(), dt=dt)


def test_oracle_matches_its_golden_fixture():
    ref = np.load(GOLDEN)
    case = _golden_case()
    o = O.Oracle(case)
    o.setup_mixing_layer()
    o.initialize()
    o.rhs(0)
    np.testing.assert_array_equal(o.array(0, O.DQ), ref["dq"])
    dt = o.step(3)
    assert dt == ref["dt"]
    np.testing.assert_array_equal(o.array(0, O.BV), ref["bv_after3"])


@pytest.mark.parametrize("scheme", [51, 52, 71, 3])
def test_rhs_is_equivariant_under_axis_permutation(scheme):
    """Swapping x and y in the state (u <-> v) swaps them in the residual
    (rho u <-> rho v rows, i <-> j indices), inviscid and viscous fluxes
    together, on a periodic box with the uniform metric.  The directions are
    accumulated in a different order, so agreement is to round-off."""
    n = 12
    case = box_case(n, scheme=scheme)
    case["reconstruction"] = 2

    def residual(swap):
        o = O.Oracle(case)
        exact_grid(o)
        g = o.ngg
        idx = np.arange(-g, n + g)
        ph = 2 * np.pi * (idx % (n - 1)) / (n - 1)
        Z, Y, X = np.meshgrid(ph, ph, ph, indexing="ij")
        if swap:
            X, Y = Y, X
        rho = 1.2 + 0.1 * np.sin(X + 2 * Y) * np.cos(Z)
        u = 50 + 20 * np.cos(Y) + 5 * np.sin(X + Z)
        v = 10 * np.sin(Z - X) + 3 * np.cos(Y)
        w = -8 * np.cos(2 * Y + Z)
        p = 1e5 + 3e3 * np.sin(X) * np.sin(2 * Y)
        if swap:
            u, v = v, u
        set_state(o, rho, u, v, w, p)
        L = O.lib()
        L.orc_apply_bc(o.ctx, 0)
        L.orc_properties(o.ctx, 0)
        o.rhs(0)
        return o.array(0, O.DQ).copy()

    a, b = residual(False), residual(True)
    b = b[[0, 2, 1, 3, 4]].transpose(0, 1, 3, 2)  # rows rho u <-> rho v, axes j <-> i
    for l in range(5):
        scale = np.max(np.abs(a[l]))
        assert np.max(np.abs(a[l] - b[l])) <= 1e-12 * scale, (l, np.max(np.abs(a[l] - b[l])) / scale)


def test_residual_is_the_rms_of_the_primitive_changes():
    """compute_residual (src/Residual.cuh:12-187): sqrt(sum(dv^2) / N) over the
    interior of every block for rho, |V|, p and T, against numpy."""
    case = cases.small_multiblock_case()
    o = O.Oracle(case)
    o.setup_mixing_layer()
    o.initialize()
    L = O.lib()
    assert L.orc_store_last_step(o.ctx) == 0
    g = o.ngg
    before = [o.array(b, O.BV)[:, g:-g, g:-g, g:-g].copy() for b in range(len(case["blocks"]))]
    o.step(2)
    acc, n = np.zeros(4), 0
    for b, old in enumerate(before):
        new = o.array(b, O.BV)[:, g:-g, g:-g, g:-g]
        vel = lambda q: np.sqrt(q[1] ** 2 + q[2] ** 2 + q[3] ** 2)
        acc += [np.sum((new[0] - old[0]) ** 2), np.sum((vel(new) - vel(old)) ** 2),
                np.sum((new[4] - old[4]) ** 2), np.sum((new[5] - old[5]) ** 2)]
        n += new[0].size
    want = np.sqrt(acc / n)
    got = o.residual()
    np.testing.assert_allclose(got, want, rtol=1e-12)
    assert np.all(want > 0)
