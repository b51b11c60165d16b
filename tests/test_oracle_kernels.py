"""Pins of the oracle's point kernels (WENO-Z5/Z7, WENO5-JS, LF splitting)
against what the mathematics fixes, independent of the oracle's own code:

* polynomial exactness of the candidate stencils (any convex weights return
  the common candidate value) -- a wrong coefficient fails it;
* the published weight formulas evaluated in 50-digit arithmetic (mpmath):
  WENO-Z of Borges, Carmona, Costa & Don (J. Comput. Phys. 227, 2008) with
  tau5 = |beta0 - beta2| and exponent 2, WENO-JS of Jiang & Shu (J. Comput.
  Phys. 126, 1996) with exponent 2, WENO-Z7 of Castro, Costa & Don (2011)
  with tau7 = |beta0 - beta3|;
* mirror symmetry of the upwind and downwind halves;
* the LF split identities F+ + F- = F and F+ - F- = J*lambda*Q, with F the
  physical flux written here from the Euler equations.
"""
import mpmath as mp
import numpy as np
import pytest

from _util import O

mp.mp.dps = 50
rng = np.random.default_rng(1234)


def _mp_weno5_side(v, eps, d, kind):
    v = [mp.mpf(float(x)) for x in v]
    q = [(2 * v[2] + 5 * v[3] - v[4]) / 6, (-v[1] + 5 * v[2] + 2 * v[3]) / 6, (2 * v[0] - 7 * v[1] + 11 * v[2]) / 6]
    b = [mp.mpf(13) / 12 * (v[2] - 2 * v[3] + v[4]) ** 2 + mp.mpf(1) / 4 * (3 * v[2] - 4 * v[3] + v[4]) ** 2,
         mp.mpf(13) / 12 * (v[1] - 2 * v[2] + v[3]) ** 2 + mp.mpf(1) / 4 * (v[1] - v[3]) ** 2,
         mp.mpf(13) / 12 * (v[0] - 2 * v[1] + v[2]) ** 2 + mp.mpf(1) / 4 * (v[0] - 4 * v[1] + 3 * v[2]) ** 2]
    e = mp.mpf(float(eps))
    if kind == "z":
        tau = abs(b[0] - b[2])
        a = [d[k] * (1 + (tau / (e + b[k])) ** 2) for k in range(3)]
    else:
        a = [d[k] / (e + b[k]) ** 2 for k in range(3)]
    return sum(a[k] * q[k] for k in range(3)) / sum(a)


def mp_weno5(vp, vm, eps, kind="z"):
    # upwind half on vp (points i-2..i+2), downwind half = mirror image on vm
    plus = _mp_weno5_side(vp, eps, [mp.mpf(3) / 10, mp.mpf(6) / 10, mp.mpf(1) / 10], kind)
    minus = _mp_weno5_side(vm[::-1], eps, [mp.mpf(3) / 10, mp.mpf(6) / 10, mp.mpf(1) / 10], kind)
    return plus + minus


def mp_weno7_side(v, eps):
    v = [mp.mpf(float(x)) for x in v]
    q = [(-3 * v[0] + 13 * v[1] - 23 * v[2] + 25 * v[3]) / 12, (v[1] - 5 * v[2] + 13 * v[3] + 3 * v[4]) / 12,
         (-v[2] + 7 * v[3] + 7 * v[4] - v[5]) / 12, (3 * v[3] + 13 * v[4] - 5 * v[5] + v[6]) / 12]
    # Balsara & Shu (2000) smoothness indicators, written through the first,
    # second and third derivative-like differences of each 4-point stencil
    def beta(s1, s2, s3):
        return s1 * s1 + mp.mpf(13) / 12 * s2 * s2 + mp.mpf(1043) / 960 * s3 * s3 + mp.mpf(1) / 12 * s1 * s3
    b = [beta((-2 * v[0] + 9 * v[1] - 18 * v[2] + 11 * v[3]) / 6, -v[0] + 4 * v[1] - 5 * v[2] + 2 * v[3],
              -v[0] + 3 * v[1] - 3 * v[2] + v[3]),
         beta((v[1] - 6 * v[2] + 3 * v[3] + 2 * v[4]) / 6, v[2] - 2 * v[3] + v[4], -v[1] + 3 * v[2] - 3 * v[3] + v[4]),
         beta((-2 * v[2] - 3 * v[3] + 6 * v[4] - v[5]) / 6, v[2] - 2 * v[3] + v[4], -v[2] + 3 * v[3] - 3 * v[4] + v[5]),
         beta((-11 * v[3] + 18 * v[4] - 9 * v[5] + 2 * v[6]) / 6, 2 * v[3] - 5 * v[4] + 4 * v[5] - v[6],
              -v[3] + 3 * v[4] - 3 * v[5] + v[6])]
    d = [mp.mpf(1) / 35, mp.mpf(12) / 35, mp.mpf(18) / 35, mp.mpf(4) / 35]
    e = mp.mpf(float(eps))
    tau = abs(b[0] - b[3])
    a = [d[k] * (1 + (tau / (e + b[k])) ** 2) for k in range(4)]
    return sum(a[k] * q[k] for k in range(4)) / sum(a)


def test_weno5_candidates_exact_for_quadratics():
    for _ in range(50):
        c0, c1, c2 = rng.normal(size=3)
        x = np.arange(-2, 3, dtype=float)
        f = c0 + c1 * x + c2 * x * x
        # every candidate reconstructs h(1/2) of the primitive-function
        # derivative, which for a quadratic is f(1/2) - f''/24
        exact = c0 + c1 * 0.5 + c2 * 0.25 - 2 * c2 / 24
        for fn in (O.weno5z, O.weno5js):
            got = fn(f, np.zeros(5), 1e-6)
            assert abs(got - exact) <= 1e-13 * max(1.0, abs(exact))


def test_weno7_candidates_exact_for_cubics():
    for _ in range(50):
        c = rng.normal(size=4)
        x = np.arange(-3, 4, dtype=float)
        f = c[0] + c[1] * x + c[2] * x ** 2 + c[3] * x ** 3
        # h(x) with (1/dx) int_{x-1/2}^{x+1/2} h = f: h = f - f''/24 (exact to cubic order)
        xh = 0.5
        exact = (c[0] + c[1] * xh + c[2] * xh ** 2 + c[3] * xh ** 3) - (2 * c[2] + 6 * c[3] * xh) / 24
        got = O.weno7z(f, np.zeros(7), 1e-6)
        assert abs(got - exact) <= 1e-12 * max(1.0, abs(exact))


def test_constant_data_returns_the_constant():
    for fn, n in ((O.weno5z, 5), (O.weno5js, 5), (O.weno7z, 7)):
        for v in (1.0, -3.25, 1e-7):
            assert fn(np.full(n, v), np.full(n, 0.0), 1e-30) == pytest.approx(v, rel=1e-15)


@pytest.mark.parametrize("kind", ["z", "js"])
def test_weno5_matches_published_weights_in_50_digits(kind):
    fn = O.weno5z if kind == "z" else O.weno5js
    for trial in range(300):
        scale = 10.0 ** rng.uniform(-8, 3)
        if trial % 3 == 0:  # a discontinuity inside the stencil
            vp = np.where(np.arange(5) < rng.integers(1, 5), 1.0, -2.0) * scale + rng.normal(size=5) * scale * 1e-3
            vm = np.where(np.arange(5) < rng.integers(1, 5), 0.5, 3.0) * scale + rng.normal(size=5) * scale * 1e-3
        else:
            vp, vm = rng.normal(size=5) * scale, rng.normal(size=5) * scale
        eps = 1e-20 * scale * scale
        ref = mp_weno5(vp, vm, eps, kind)
        got = fn(vp, vm, eps)
        mag = float(max(np.max(np.abs(vp)), np.max(np.abs(vm))))
        assert abs(got - float(ref)) <= 2e-14 * mag, (trial, got, ref)


def test_weno7_matches_published_weights_in_50_digits():
    for trial in range(200):
        scale = 10.0 ** rng.uniform(-6, 2)
        vp, vm = rng.normal(size=7) * scale, rng.normal(size=7) * scale
        eps = 1e-20 * scale * scale
        ref = mp_weno7_side(vp, eps) + mp_weno7_side(vm[::-1], eps)
        got = O.weno7z(vp, vm, eps)
        mag = float(max(np.max(np.abs(vp)), np.max(np.abs(vm))))
        assert abs(got - float(ref)) <= 5e-14 * mag, (trial, got, ref)


def test_upwind_and_downwind_halves_are_mirror_images():
    for _ in range(100):
        v = rng.normal(size=5)
        a = O.weno5z(v, np.zeros(5), 1e-12)          # upwind half only
        b = O.weno5z(np.zeros(5), v[::-1], 1e-12)    # downwind half on the mirrored data
        assert a == pytest.approx(b, rel=1e-14, abs=1e-15)


def test_lf_split_identities():
    gamma = 1.4
    for _ in range(200):
        rho = rng.uniform(0.1, 5)
        u, v, w = rng.normal(size=3) * 300
        p = rng.uniform(1e3, 1e6)
        E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w)
        c = np.sqrt(gamma * p / rho)
        k = rng.normal(size=3) * 1e4
        jac = rng.uniform(1e-12, 1e-6)
        q = np.array([rho, rho * u, rho * v, rho * w, E, p, c])
        fp, fm = O.lf_split(q, k, jac)
        U = u * k[0] + v * k[1] + w * k[2]
        F = jac * np.array([rho * U, rho * u * U + p * k[0], rho * v * U + p * k[1], rho * w * U + p * k[2],
                            (E + p) * U])
        lam = abs(U) + c * np.linalg.norm(k)
        scale = np.max(np.abs(F)) + jac * lam * np.max(np.abs(q[:5]))
        np.testing.assert_allclose(fp + fm, F, rtol=0, atol=1e-14 * scale)
        np.testing.assert_allclose(fp - fm, jac * lam * q[:5], rtol=0, atol=1e-14 * scale)
        # the upwind half carries no left-going characteristic: F+ >= 0 for rho
        assert fp[0] >= 0 >= fm[0]


@pytest.mark.parametrize("order", [5, 7])
def test_flux_difference_converges_at_the_design_order(order):
    """Finite-difference WENO: (h(i+1/2) - h(i-1/2)) / dx with the upwind half
    on point values of f = sin(x) (downwind half zero) approximates f'(x_i);
    in smooth data the WENO-Z weights reach the optimal ones fast enough that
    halving dx divides the error by about 2^order."""
    fn = O.weno5z if order == 5 else O.weno7z
    w = order
    errs = []
    for n in (16, 32, 64):
        dx = 2 * np.pi / n
        x = np.arange(-w, n + w) * dx + 0.3
        f = np.sin(x)
        zero = np.zeros(w)

        def h(i):  # face i+1/2 of point index i (offset w in the padded array)
            c = i + w
            return fn(f[c - w // 2:c + w // 2 + 1], zero, 1e-40)

        d = np.array([(h(i) - h(i - 1)) / dx for i in range(n)])
        errs.append(np.max(np.abs(d - np.cos(x[w:n + w]))))
    rates = [np.log2(errs[k] / errs[k + 1]) for k in range(2)]
    assert rates[1] > order - 0.5, (errs, rates)
