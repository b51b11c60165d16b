"""Pins of the oracle's farfield boundary (label 4) for air (CPU only).

FarField constructor (src/BoundCond.cpp:586-700): the three input
combinations against the ideal-gas closed forms.

apply_farfield<Air> (src/BoundCond.cuh:755-1022), on a box whose interior
normal velocity sweeps through every regime along x and whose face metric is
tilted, checked point by point through what the characteristic theory fixes:
  supersonic inflow   face point and ghosts = the far-field state;
  supersonic outflow  face point unchanged, ghosts = the face point;
  subsonic            the boundary state carries the outgoing invariant
                      u_n + 2c/(gamma-1) of the face point and the incoming
                      u_n - 2c/(gamma-1) of the far field, the far-field
                      entropy and tangential velocity on inflow, the face
                      point's on outflow, T from the gas law; ghosts = face.
A far field equal to the interior state is reproduced.
"""
import numpy as np
import pytest

from _util import O, cases
from test_oracle_rhs import box_case, exact_grid, set_state

GAMMA, RU, MW = 1.4, cases.R_U, cases.MW_AIR
G5 = 2 / (GAMMA - 1)


def test_farfield_state_mach_pressure_temperature():
    d = np.array([0.6, 0.8, 0.0])
    f = O.farfield_state_air(mach=0.7, pressure=9.0e4, temperature=280.0, u=d[0], v=d[1], w=d[2])
    c = np.sqrt(GAMMA * RU / MW * 280.0)
    rho = 9.0e4 * MW / (RU * 280.0)
    assert np.isclose(f.acoustic_speed, c, rtol=1e-15)
    assert np.isclose(f.density, rho, rtol=1e-15)
    np.testing.assert_allclose([f.u, f.v, f.w], 0.7 * c * d, rtol=1e-15)
    assert np.isclose(f.entropy, 9.0e4 / rho ** GAMMA, rtol=1e-14)
    assert (f.pressure, f.temperature, f.specific_heat_ratio, f.mw) == (9.0e4, 280.0, GAMMA, MW)


def test_farfield_state_density_velocity_pressure():
    f = O.farfield_state_air(density=1.1, velocity=250.0, pressure=1.0e5)
    T = 1.0e5 * MW / (1.1 * RU)
    assert np.isclose(f.temperature, T, rtol=1e-15)
    assert np.isclose(f.u, 250.0, rtol=1e-15) and f.v == 0 and f.w == 0
    assert np.isclose(f.density, 1.1, rtol=0)


def test_farfield_state_mach_temperature_reynolds():
    f = O.farfield_state_air(mach=2.0, temperature=220.0, reynolds_number=3.0e6)
    c = np.sqrt(GAMMA * RU / MW * 220.0)
    mu = O.lib().orc_sutherland(220.0)
    rho = mu * 3.0e6 / (2.0 * c)
    assert np.isclose(f.density, rho, rtol=1e-14)
    assert np.isclose(f.pressure, rho * 220.0 * RU / MW, rtol=1e-14)


def _case(n, lo, hi):
    c = box_case(n, faces={0: cases.BC_PERIODIC, 1: cases.BC_FARFIELD, 2: cases.BC_PERIODIC})
    for f in c["boundaries"]:
        if f["type"] == cases.BC_FARFIELD:
            f["farfield"] = hi if f["side"] > 0 else lo
    return c


LO = dict(mach=0.6, pressure=1.1e5, temperature=320.0, u=0.6, v=0.0, w=0.8)
HI = dict(mach=0.5, pressure=9.0e4, temperature=280.0, u=0.6, v=0.8, w=0.0)


def _apply(n=11, tilt=(0.3, -0.5)):
    o = O.Oracle(_case(n, LO, HI))
    g = o.ngg
    exact_grid(o)
    met = o.array(0, O.METRIC)
    met[3][...] = tilt[0]
    met[5][...] = tilt[1]
    # interior normal speed sweeping -2.2c .. 2.2c along x; tangential parts
    rho, p = 1.15, 1.0e5
    c = np.sqrt(GAMMA * p / rho)
    shp = o.shape(0, O.BV)
    v = np.broadcast_to(np.linspace(-2.2 * c, 2.2 * c, shp[2]), shp).copy()
    set_state(o, rho, 120.0, v, -40.0, p)
    pre = o.array(0, O.BV).copy()
    assert O.lib().orc_apply_bc(o.ctx, 0) == 0
    return o, g, pre, c


def _check_face(o, g, pre, side, ffspec):
    ff = O.farfield_state_air(**ffspec)
    uinf = np.array([ff.u, ff.v, ff.w])
    bv, cv, met = o.array(0, O.BV), o.array(0, O.CV), o.array(0, O.METRIC)
    n = o.case["blocks"][0]["my"]
    jf = g if side < 0 else g + n - 1
    counts = dict(sup_in=0, sup_out=0, sub_in=0, sub_out=0)
    for k in range(g, g + n):
        for i in range(g, g + n):
            m = met[3:6, k, jf, i]
            nh = side * m / np.linalg.norm(m)
            q0 = pre[:, k, jf, i]
            a = np.sqrt(GAMMA * q0[4] / q0[0])
            un0 = nh @ q0[1:4]
            M = un0 / a
            qb = bv[:, k, jf, i]
            ghosts = [bv[:, k, jf + side * gg, i] for gg in range(1, g + 1)]
            for q in ghosts:
                np.testing.assert_array_equal(q, qb)
            # cv of the face point and the ghosts follow from bv
            for gg in range(0, g + 1):
                jj = jf + side * gg
                q = bv[:, k, jj, i]
                E = q[4] / (GAMMA - 1) + 0.5 * q[0] * (q[1:4] @ q[1:4])
                np.testing.assert_allclose(cv[:, k, jj, i], [q[0], *(q[0] * q[1:4]), E], rtol=1e-14)
            if M <= -1:
                counts["sup_in"] += 1
                np.testing.assert_array_equal(qb, [ff.density, ff.u, ff.v, ff.w, ff.pressure, ff.temperature])
                continue
            if M >= 1:
                counts["sup_out"] += 1
                np.testing.assert_array_equal(qb, q0)
                continue
            cb = np.sqrt(GAMMA * qb[4] / qb[0])
            unb = nh @ qb[1:4]
            scale = a + abs(un0) + ff.acoustic_speed
            assert abs((unb + G5 * cb) - (un0 + G5 * a)) <= 1e-13 * G5 * scale
            assert abs((unb - G5 * cb) - (nh @ uinf - G5 * ff.acoustic_speed)) <= 1e-13 * G5 * scale
            tb = qb[1:4] - unb * nh
            if M <= 0:
                counts["sub_in"] += 1
                assert np.isclose(qb[4] / qb[0] ** GAMMA, ff.entropy, rtol=1e-13)
                np.testing.assert_allclose(tb, uinf - (nh @ uinf) * nh, atol=1e-12 * scale)
            else:
                counts["sub_out"] += 1
                assert np.isclose(qb[4] / qb[0] ** GAMMA, q0[4] / q0[0] ** GAMMA, rtol=1e-13)
                np.testing.assert_allclose(tb, q0[1:4] - un0 * nh, atol=1e-12 * scale)
            assert np.isclose(qb[5], qb[4] * MW / (qb[0] * RU), rtol=1e-15)
    return counts


@pytest.mark.parametrize("side", [-1, 1])
def test_farfield_regimes(side):
    o, g, pre, _ = _apply()
    counts = _check_face(o, g, pre, side, HI if side > 0 else LO)
    assert all(v > 0 for v in counts.values()), counts


def test_farfield_reproduces_a_matching_far_field():
    """Interior state = far-field state (subsonic both ways): the boundary
    keeps it."""
    spec = dict(mach=0.3, pressure=1.0e5, temperature=300.0, u=0.0, v=1.0, w=0.0)
    o = O.Oracle(_case(9, spec, spec))
    exact_grid(o)
    ff = O.farfield_state_air(**spec)
    set_state(o, ff.density, ff.u, ff.v, ff.w, ff.pressure)
    assert O.lib().orc_apply_bc(o.ctx, 0) == 0
    bv = o.array(0, O.BV)
    want = [ff.density, ff.u, ff.v, ff.w, ff.pressure, ff.temperature]
    for l in range(6):
        np.testing.assert_allclose(bv[l], want[l], rtol=1e-13, atol=1e-13 * abs(ff.v))
