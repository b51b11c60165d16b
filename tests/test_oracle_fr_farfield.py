"""Pins of the mixture oracle's farfield boundary (label 4) (CPU only).

FarField constructor with species (src/BoundCond.cpp:586-700): gamma =
cp/cv of the composition from an independent NASA-7 evaluation, the sound
speed, density and entropy of the ideal-gas mixture; the (mach, temperature,
reynolds_number) combination against the mixture-averaged viscosity of the
transport routine.

apply_farfield<FR> (src/BoundCond.cuh:755-1022) on a box whose normal
velocity sweeps every regime, interior and far-field compositions differing:
supersonic inflow = the far field (state and mass fractions); supersonic
outflow = copy of the face point; subsonic: the normal velocity is the mean
of the outgoing invariant (interior gamma) and the incoming one (far-field
gamma); the gamma g of the written state, recovered from
c = sqrt(g p/rho) = (g - 1)(R+ - R-)/4, is within 1e-4 of cp/cv at the
boundary temperature and composition (the iteration's stopping rule), the
entropy p/rho^g is the far field's (inflow) or the face point's (outflow),
tangential velocity and composition come from the same side, T follows the
gas law; ghosts = face.
"""
import numpy as np
import pytest
from scipy.optimize import brentq

from _util import pkg
import mech
import oracle_fr as OF

MX = pkg.mixture
R_U = 8314.462618


@pytest.fixture(scope="module")
def T():
    return mech.load(*MX.H2_AIR_FILES)


def _cp(T, t):
    """NASA-7 cp_i (J/kg/K), evaluated at max(t, t_low)."""
    ns = T.n_spec
    cp = np.zeros(ns)
    for i in range(ns):
        tt = max(t, T.t_low[i])
        a = T.low_coeff[i] if tt < T.t_mid[i] else T.high_coeff[i]
        cp[i] = (a[0] + a[1] * tt + a[2] * tt ** 2 + a[3] * tt ** 3 + a[4] * tt ** 4) * R_U / T.mw[i]
    return cp


def _gamma(T, t, y):
    mw = np.array(T.mw[:T.n_spec])
    cp = _cp(T, t)
    return (y @ cp) / (y @ (cp - R_U / mw)), 1.0 / np.sum(y / mw)


Y_AIR = np.array([0, 0, 0, 0, 0, 0, 0.232, 0, 0.768])
Y_FUEL = np.array([0, 0.05, 0, 0, 0, 0, 0, 0, 0.95])


def test_mixture_farfield_state(T):
    f = OF.farfield_state(T, Y_AIR, mach=0.8, pressure=9.0e4, temperature=1500.0, u=0.6, v=0.8)
    g, mw = _gamma(T, 1500.0, Y_AIR)
    c = np.sqrt(g * R_U / mw * 1500.0)
    rho = 9.0e4 * mw / (R_U * 1500.0)
    assert np.isclose(f.specific_heat_ratio, g, rtol=1e-14)
    assert np.isclose(f.mw, mw, rtol=1e-15)
    assert np.isclose(f.acoustic_speed, c, rtol=1e-14)
    assert np.isclose(f.density, rho, rtol=1e-14)
    np.testing.assert_allclose([f.u, f.v, f.w], [0.48 * c, 0.64 * c, 0.0], rtol=1e-14)
    assert np.isclose(f.entropy, 9.0e4 / rho ** g, rtol=1e-13)
    np.testing.assert_array_equal(np.array(f.sv[:9]), Y_AIR)


def test_mixture_farfield_state_from_reynolds_number(T):
    f = OF.farfield_state(T, Y_FUEL, mach=1.5, temperature=400.0, reynolds_number=2.0e6)
    mu, _, _ = OF.transport(T, 400.0, f.pressure, f.density, Y_FUEL)
    assert np.isclose(f.density * np.hypot(f.u, f.v) / 2.0e6, mu, rtol=1e-13)
    assert np.isclose(f.pressure, f.density * 400.0 * R_U / f.mw, rtol=1e-14)


def _case(T, n, lo, hi):
    c = pkg.cases.reacting_single_block(n, n, n, T, scheme=52, viscous_order=2, reaction=0)
    c["boundaries"] = [dict(block=0, axis=a, side=s, type=8) for a in (0, 2) for s in (-1, 1)]
    for s, (spec, y) in ((-1, lo), (1, hi)):
        c["boundaries"].insert(0, dict(block=0, axis=1, side=s, type=4, farfield=spec, farfield_y=list(y)))
    return c


LO = (dict(mach=0.6, pressure=1.1e5, temperature=1500.0, u=0.6, v=0.0, w=0.8), Y_AIR)
HI = (dict(mach=0.5, pressure=9.0e4, temperature=300.0, u=0.6, v=0.8, w=0.0), Y_FUEL)
Y_IN = np.array([0.01, 0.02, 0.05, 0.001, 0.002, 0.01, 0.2, 0.02, 0.0])


def _apply(T, n=9):
    y_in = Y_IN.copy()
    y_in[8] = 1 - y_in[:8].sum()
    o = OF.FrOracle(_case(T, n, LO, HI), T, n_threads=4)
    jac, met = o.array(0, OF.JAC), o.array(0, OF.METRIC)
    jac[...] = 0.125
    met[...] = 0.0
    met[0], met[4], met[8] = 2.0, 2.0, 2.0
    met[3], met[5] = 0.3, -0.5  # tilted face normal
    bv, sv = o.array(0, OF.BV), o.array(0, OF.SV)
    sv[...] = y_in[:, None, None, None]
    g0, mw0 = _gamma(T, 900.0, y_in)
    rho, p = 0.6, 0.6 * R_U / mw0 * 900.0
    c = np.sqrt(g0 * p / rho)
    shp = bv.shape[1:]
    bv[0], bv[1], bv[3], bv[4], bv[5] = rho, 120.0, -40.0, p, 900.0
    bv[2] = np.broadcast_to(np.linspace(-2.2 * c, 2.2 * c, shp[2]), shp)
    o.L.orf_cv_from_bv(o.ctx, 0)
    o.L.orf_properties(o.ctx, 0)
    pre_bv, pre_sv, pre_gam = bv.copy(), sv.copy(), o.array(0, OF.GAMMA)[0].copy()
    assert o.L.orf_apply_bc(o.ctx, 0) == 0
    return o, pre_bv, pre_sv, pre_gam


@pytest.mark.parametrize("side", [-1, 1])
def test_mixture_farfield_regimes(T, side):
    o, pre_bv, pre_sv, pre_gam = _apply(T)
    spec, y_ff = HI if side > 0 else LO
    ff = OF.farfield_state(T, y_ff, **spec)
    uinf = np.array([ff.u, ff.v, ff.w])
    g = o.ngg
    n = o.case["blocks"][0]["my"]
    bv, sv, met = o.array(0, OF.BV), o.array(0, OF.SV), o.array(0, OF.METRIC)
    jf = g if side < 0 else g + n - 1
    counts = dict(sup_in=0, sup_out=0, sub_in=0, sub_out=0)
    for k in range(g, g + n):
        for i in range(g, g + n):
            m = met[3:6, k, jf, i]
            nh = side * m / np.linalg.norm(m)
            q0, y0, g_b = pre_bv[:, k, jf, i], pre_sv[:, k, jf, i], pre_gam[k, jf, i]
            a = np.sqrt(g_b * q0[4] / q0[0])
            un0 = nh @ q0[1:4]
            M = un0 / a
            qb, yb = bv[:, k, jf, i], sv[:, k, jf, i]
            for gg in range(1, g + 1):
                np.testing.assert_array_equal(bv[:, k, jf + side * gg, i], qb)
                np.testing.assert_array_equal(sv[:, k, jf + side * gg, i], yb)
            if M <= -1:
                counts["sup_in"] += 1
                np.testing.assert_array_equal(qb, [ff.density, ff.u, ff.v, ff.w, ff.pressure, ff.temperature])
                np.testing.assert_array_equal(yb, y_ff)
                continue
            if M >= 1:
                counts["sup_out"] += 1
                np.testing.assert_array_equal(qb, q0)
                np.testing.assert_array_equal(yb, y0)
                continue
            rp = un0 + 2 * a / (g_b - 1)
            rm = nh @ uinf - 2 * ff.acoustic_speed / (ff.specific_heat_ratio - 1)
            unb = nh @ qb[1:4]
            assert abs(unb - 0.5 * (rp + rm)) <= 1e-12 * (abs(rp) + abs(rm))
            gk = brentq(lambda x: 0.25 * (x - 1) * (rp - rm) - np.sqrt(x * qb[4] / qb[0]), 1.01, 2.5, xtol=1e-15)
            inflow = M < 0
            y_side = y_ff if inflow else y0
            np.testing.assert_array_equal(yb, y_side)
            g_T, mw = _gamma(T, qb[5], y_side)
            if inflow:
                mw = ff.mw
            assert abs(1 - g_T / gk) <= 1.05e-4, (g_T, gk)
            s_ref = ff.entropy if M <= 0 else q0[4] / q0[0] ** gk
            assert np.isclose(qb[4] / qb[0] ** gk, s_ref, rtol=1e-10)
            assert np.isclose(qb[5], qb[4] * mw / (qb[0] * R_U), rtol=1e-14)
            tb = qb[1:4] - unb * nh
            want = (uinf - (nh @ uinf) * nh) if inflow else (q0[1:4] - un0 * nh)
            np.testing.assert_allclose(tb, want, atol=1e-11 * (abs(rp) + abs(rm)))
            counts["sub_in" if inflow else "sub_out"] += 1
    assert all(v > 0 for v in counts.values()), counts
