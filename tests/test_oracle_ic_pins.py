"""Closed-form pins of the mixing-layer stream states, the initial profile
and the outflow ghost fill, on the oracle and on the product's host helpers.

What the mathematics fixes (get_mixing_layer_info, src/MixingLayer.cu:9-190;
initialize_mixing_layer_with_info, src/Initialize.cu:259-325; apply_outflow,
src/BoundCond.cuh:205-233):

* the convective Mach number of the two streams, (u1 - u2) / (c1 + c2), is the
  requested Mc (upper stream faster; (u2 - u1) / (c1 + c2) otherwise);
* each stream obeys the ideal-gas law p = rho R T, and its speed of sound is
  sqrt(gamma R T) with gamma = cp / (cp - R); for mixtures cp is checked
  against the derivative of the NASA-7 enthalpy polynomial (a different
  formula from the cp polynomial the code evaluates), read by the
  independent Python reader;
* u(y) = (u1 + u2)/2 + (u1 - u2)/2 tanh(2 y / delta_omega), so u(0) is the
  mean of the two streams and u is odd about it; every other primitive is
  the upper stream's where y >= 0 and the lower stream's below;
* the seeded perturbation is bounded by its amplitude, has zero mean and the
  variance a^2/3 of U(-a, a), independently per component;
* outflow ghost layer g of a face equals the face node (zero gradient), and
  the conservative variables of each ghost are those of its primitives.

A planted mistake (a sign in the lower Mach number, a swapped stream, the
ghost copied from the second layer, a cp of the wrong coefficient set) fails
one of these.
"""
import ctypes as C
import math

import numpy as np
import pytest

from _util import O, cases, pkg

import mech
import oracle_fr as OF

R_U = cases.R_U


def _air_streams_oracle(ma_c, ma_u, t_u, t_l, p_u, p_l, faster):
    st = O._Streams()
    O.lib().orc_mixing_layer_streams(ma_c, ma_u, t_u, t_l, p_u, p_l, int(faster), C.byref(st))
    return list(st.upper), list(st.lower)


AIR_CASES = [(0.4, 0.4, 300.0, 300.0, 101325.0, 101325.0, True),
             (0.5, 0.9, 300.0, 900.0, 101325.0, 5.0e4, True),
             (0.3, 0.2, 1200.0, 400.0, 2.0e5, 2.0e5, False),
             (0.7, 1.5, 500.0, 250.0, 8.0e4, 1.2e5, False)]


@pytest.mark.parametrize("args", AIR_CASES)
def test_air_streams_closed_forms(args):
    ma_c, ma_u, t_u, t_l, p_u, p_l, faster = args
    py = cases.air_streams(ma_c=ma_c, ma_upper=ma_u, t_upper=t_u, t_lower=t_l, p_upper=p_u, p_lower=p_l,
                           upper_faster=faster)
    orc = _air_streams_oracle(*args)
    r_air = R_U / cases.MW_AIR
    for up, lo in ((py["upper"], py["lower"]), orc):
        c_u = math.sqrt(1.4 * r_air * t_u)
        c_l = math.sqrt(1.4 * r_air * t_l)
        # ideal-gas law and the given p, T
        assert up[4] == p_u and lo[4] == p_l and up[5] == t_u and lo[5] == t_l
        assert abs(up[0] * r_air * t_u - p_u) <= 1e-14 * p_u
        assert abs(lo[0] * r_air * t_l - p_l) <= 1e-14 * p_l
        # the given upper Mach number and the convective Mach number
        assert abs(up[1] - ma_u * c_u) <= 1e-14 * abs(up[1])
        mc = (up[1] - lo[1]) / (c_u + c_l) if faster else (lo[1] - up[1]) / (c_u + c_l)
        assert abs(mc - ma_c) <= 1e-13, mc
        assert up[2] == up[3] == lo[2] == lo[3] == 0.0
    # the package's states and the oracle's restatement agree
    np.testing.assert_allclose(py["upper"], orc[0], rtol=1e-15)
    np.testing.assert_allclose(py["lower"], orc[1], rtol=1e-15)


def _cp_from_enthalpy(T, t, i, h=1e-3):
    """cp of species i [J/(kg K)] as the central difference of the NASA-7
    enthalpy polynomial h/R = a1 T + a2 T^2/2 + ... + a5 T^5/5 + a6."""
    def enth(tt):
        c = T.low_coeff[i] if tt < T.t_mid[i] else T.high_coeff[i]
        return (c[0] * tt + c[1] * tt ** 2 / 2 + c[2] * tt ** 3 / 3 + c[3] * tt ** 4 / 4 + c[4] * tt ** 5 / 5
                + c[5]) * R_U / T.mw[i]
    return (enth(t + h) - enth(t - h)) / (2 * h)


@pytest.fixture(scope="module")
def h2_tables():
    return mech.load(*pkg.mixture.H2_AIR_FILES)


@pytest.mark.parametrize("ma_c", [0.5, 0.3])
def test_mixture_streams_closed_forms(h2_tables, ma_c):
    T = h2_tables
    names = mech.species_names(T)
    fuel = np.zeros(T.n_spec)
    fuel[names.index("H2")], fuel[names.index("N2")] = 0.05, 0.95
    ox = np.zeros(T.n_spec)
    ox[names.index("O2")], ox[names.index("N2")] = 0.232, 0.768
    t_u, t_l, p = 300.0, 1500.0, 101325.0
    pkg_tables = pkg.mixture.load_mechanism(*pkg.mixture.H2_AIR_FILES)
    py = cases.mixture_streams(pkg_tables, fuel, ox, t_u, t_l, p, p, ma_c=ma_c)
    orc = OF.mixture_streams(T, dict(y_upper=fuel, y_lower=ox, t_upper=t_u, t_lower=t_l, p_upper=p, p_lower=p,
                                     ma_c=ma_c))
    c = {}
    for side, y, t in (("upper", fuel, t_u), ("lower", ox, t_l)):
        r_mix = R_U * sum(y[i] / T.mw[i] for i in range(T.n_spec))
        cp = sum(y[i] * _cp_from_enthalpy(T, t, i) for i in range(T.n_spec))
        c[side] = math.sqrt(cp / (cp - r_mix) * r_mix * t)
        for st in (py, orc):
            s = st[side]
            assert abs(s[0] * r_mix * t - p) <= 1e-13 * p
            assert s[4] == p and s[5] == t and s[2] == s[3] == 0.0
    for st in (py, orc):
        # speed of sound from gamma = cp / (cp - R): the difference quotient
        # of the enthalpy is good to ~1e-9
        mc = (st["upper"][1] - st["lower"][1]) / (c["upper"] + c["lower"])
        assert abs(mc - ma_c) <= 1e-8, mc
        assert abs(st["upper"][1] - ma_c * c["upper"]) <= 1e-8 * abs(st["upper"][1])
    np.testing.assert_allclose(py["upper"], orc["upper"], rtol=1e-14)
    np.testing.assert_allclose(py["lower"], orc["lower"], rtol=1e-14)


def _ic_case(n=9, amplitude=0.0):
    c = cases.single_block_case(n)
    c["amplitude"] = amplitude
    # put a node exactly on y = 0: y0 + j dy with y0 = -4 dy
    blk = c["blocks"][0]
    blk["y0"] = -4.0 * c["dy"]
    return c


def _check_profile(bv, case, g):
    st = case["streams"]
    up, lo = st["upper"], st["lower"]
    blk = case["blocks"][0]
    y = blk["y0"] + np.arange(-g, blk["my"] + g) * case["dy"]
    d = case["delta_omega"]
    u_want = 0.5 * (up[1] + lo[1]) + 0.5 * (up[1] - lo[1]) * np.tanh(2 * y / d)
    for j, yy in enumerate(y):
        np.testing.assert_allclose(bv[1, :, j, :], u_want[j], rtol=4e-16, atol=4e-16 * abs(up[1]))
        ref = up if yy >= 0 else lo
        for l in (0, 2, 3, 4, 5):
            assert np.all(bv[l, :, j, :] == ref[l]), (j, l)
    j0 = int(np.argmin(np.abs(y)))
    assert y[j0] == 0.0
    # u at y = 0 is the mean of the streams, and u - mean is odd in y
    np.testing.assert_allclose(bv[1, :, j0, :], 0.5 * (up[1] + lo[1]), rtol=1e-15)
    dev = bv[1, 0, :, 0] - 0.5 * (up[1] + lo[1])
    k = min(j0, len(y) - 1 - j0)
    np.testing.assert_allclose(dev[j0 - k:j0][::-1], -dev[j0 + 1:j0 + 1 + k], rtol=1e-13)


def test_mixing_layer_profile_oracle():
    case = _ic_case()
    o = O.Oracle(case)
    o.setup_mixing_layer()
    _check_profile(o.array(0, O.BV), case, o.ngg)


def _product_bv(case, g):
    """The product's host helper ostc_util_mixing_layer_bv (no device work)."""
    c, blk = case, case["blocks"][0]
    bv = np.zeros((6, blk["mz"] + 2 * g, blk["my"] + 2 * g, blk["mx"] + 2 * g))
    up = np.ascontiguousarray(c["streams"]["upper"], dtype=np.float64)
    lo = np.ascontiguousarray(c["streams"]["lower"], dtype=np.float64)
    per = c["periods"]
    dp = C.POINTER(C.c_double)
    r = pkg.lib().ostc_util_mixing_layer_bv(blk["mx"], blk["my"], blk["mz"], g, blk["y0"], c["dy"],
                                            up.ctypes.data_as(dp), lo.ctypes.data_as(dp), c["delta_omega"],
                                            blk["gi0"], blk["gj0"], blk["gk0"], per[0], per[1], per[2], c["seed"],
                                            c["amplitude"], bv.ctypes.data_as(dp))
    assert r == 0
    return bv


def test_mixing_layer_profile_product_helper():
    case = _ic_case()
    _check_profile(_product_bv(case, 3), case, 3)


@pytest.mark.parametrize("which", ["oracle", "product"])
def test_seeded_perturbation_distribution(which):
    a = 2.5
    n = 40
    case = _ic_case(n, amplitude=a)
    base = _ic_case(n, amplitude=0.0)
    if which == "oracle":
        o, o0 = O.Oracle(case), O.Oracle(base)
        o.setup_mixing_layer()
        o0.setup_mixing_layer()
        d = o.array(0, O.BV)[1:4] - o0.array(0, O.BV)[1:4]
        rest = (o.array(0, O.BV)[[0, 4, 5]], o0.array(0, O.BV)[[0, 4, 5]])
    else:
        b1, b0 = _product_bv(case, 3), _product_bv(base, 3)
        d = b1[1:4] - b0[1:4]
        rest = (b1[[0, 4, 5]], b0[[0, 4, 5]])
    assert np.array_equal(*rest)                  # only velocities are perturbed
    assert np.all(np.abs(d) <= a * (1 + 1e-12))
    m = d.reshape(3, -1)
    nn = m.shape[1]
    assert np.all(np.abs(m.mean(axis=1)) <= 4 * a / math.sqrt(3 * nn))
    np.testing.assert_allclose(m.var(axis=1), a * a / 3, rtol=0.05)
    cc = np.corrcoef(m)
    assert np.all(np.abs(cc[np.triu_indices(3, 1)]) < 0.05)


def test_product_and_oracle_perturbations_agree():
    case = _ic_case(12, amplitude=1.0)
    o = O.Oracle(case)
    o.setup_mixing_layer()
    np.testing.assert_array_equal(_product_bv(case, o.ngg)[[0, 2, 3, 4, 5]], o.array(0, O.BV)[[0, 2, 3, 4, 5]])
    np.testing.assert_allclose(_product_bv(case, o.ngg)[1], o.array(0, O.BV)[1], rtol=1e-15, atol=1e-13)


def test_outflow_ghosts_copy_the_face_node():
    n = 8
    case = cases.single_block_case(n)
    o = O.Oracle(case)          # periodic x/z, outflow y
    o.setup_mixing_layer()
    g = o.ngg
    rng = np.random.default_rng(11)
    bv = o.array(0, O.BV)
    shp = bv.shape[1:]
    bv[0] = 1.0 + 0.2 * rng.random(shp)
    bv[1:4] = 50.0 * (rng.random((3, *shp)) - 0.5)
    bv[4] = 1e5 * (1.0 + 0.1 * rng.random(shp))
    bv[5] = bv[4] * cases.MW_AIR / (bv[0] * R_U)
    before = bv.copy()
    assert O.lib().orc_apply_bc(o.ctx, 0) == 0
    cv = o.array(0, O.CV)
    inner = slice(g, g + n)      # interior i and k: untouched by the periodic faces
    for gg in range(1, g + 1):
        for lay, face in ((g - gg, g), (g + n - 1 + gg, g + n - 1)):
            np.testing.assert_array_equal(bv[:, inner, lay, inner], before[:, inner, face, inner])
            rho, u, v, w, p = (bv[l, inner, lay, inner] for l in range(5))
            want = [rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)]
            for l in range(5):
                np.testing.assert_allclose(cv[l, inner, lay, inner], want[l], rtol=1e-14)
    # the interior is not touched
    np.testing.assert_array_equal(bv[:, inner, g:g + n, inner], before[:, inner, g:g + n, inner])
