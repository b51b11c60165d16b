"""The oracle's operation-counting build (-DORC_COUNT, oracle/count.h).

Pins:
* the counting build computes the same values as the double build, bit for
  bit, after whole SSP-RK3 steps;
* the counts of the point routines equal hand counts of the reference's
  formulas: one WENO-Z5 face value (src/WENO.cu:1582-1617, both sides) and
  one Lax-Friedrichs split of the five air fluxes (src/WENO.cu:697-720);
* the counts of the inviscid and viscous operators on blocks of different
  sizes equal the pencil structure of the reference's launches: per
  direction, (m + 2 ngg) split-flux points, (m + 1) faces and m cells per
  pencil, with the per-point, per-face and per-cell counts derived by hand
  below.  A routine evaluated twice, a face loop of the wrong length or a
  lost term changes one of these numbers.

The hand counts (additions include subtractions; |x| and negation are free):
* WENO-Z5, one side: candidates 10 mul 6 add; indicators 19 mul 14 add;
  weights 7 mul 7 add 3 div; the weighted sum 3 mul 4 add 1 div.  Both sides
  and their sum: 78 mul, 63 add, 8 div.
* LF split: Uk 3 mul 2 add 1 div; c|grad k| 4 mul 2 add 1 sqrt; lambda 1 add;
  0.5 J once 1 mul; F+ 18 mul 9 add; F- 18 mul 9 add: 44 mul 23 add 1 div 1 sqrt.
* inviscid point: the LF split plus c = sqrt(gamma R T): +1 mul +1 sqrt.
* inviscid face: |m1 J1 + m2 J2|^2 9 mul 5 add, the epsilon scalings 5 mul,
  five WENO-Z5 values.
* viscous face (src/ViscousScheme.cuh:23-318): face metric 9 mul 9 add;
  differences 8 mul 28 add; gradients of u, v, w, T 36 mul 24 add; mu and the
  laminar sum mu + mu_t 1 mul 2 add; normal stresses 12 mul 6 add 3 div;
  shear stresses 3 mul 3 add; k J 9 mul 3 add; momentum fluxes 9 mul 6 add;
  face velocity 3 mul 3 add; conductivity 1 mul 1 add 1 div; energy flux
  7 mul 5 add: 98 mul 90 add 4 div.
* each cell: dq -/+= f(i+1/2) - f(i-1/2) for 5 variables: 10 add.
"""
import ctypes as C

import numpy as np
import pytest

from _util import O, cases

DP = C.POINTER(C.c_double)
KEYS = ("add", "mul", "div", "sqrt", "trans")


def v(**kw):
    return np.array([kw.get(k, 0) for k in KEYS], dtype=np.int64)


WENO5Z = v(add=63, mul=78, div=8)
LF5 = v(add=23, mul=44, div=1, sqrt=1)
INV_POINT = LF5 + v(mul=1, sqrt=1)
INV_FACE = v(mul=14, add=5) + 5 * WENO5Z
VISC_FACE = v(mul=98, add=90, div=4)
CELL = v(add=10)


def counts(L):
    out = (C.c_ulonglong * 5)()
    L.orc_op_counts(out)
    return np.array(list(out), dtype=np.int64)


def test_count_build_matches_double_build():
    case = cases.small_multiblock_case()
    a, b = O.Oracle(case), O.Oracle(case, tangent="count")
    for o in (a, b):
        o.setup_mixing_layer()
        o.initialize()
    dt_a, dt_b = a.step(3), b.step(3)
    assert dt_a == dt_b
    for blk in range(len(case["blocks"])):
        for w in (O.CV, O.BV, O.MUL):
            np.testing.assert_array_equal(a.array(blk, w), b.array(blk, w))


def test_point_routine_counts():
    L = O.lib("count")
    rng = np.random.default_rng(1)
    vp, vm = rng.random(5), rng.random(5)
    L.orc_op_counts_reset()
    L.orc_weno5z(vp.ctypes.data_as(DP), vm.ctypes.data_as(DP), 1e-6)
    np.testing.assert_array_equal(counts(L), WENO5Z)
    q = np.array([1.2, 100.0, 20.0, -5.0, 2.5e5, 1e5, 340.0])
    k = np.array([0.5, 0.1, 0.2])
    fp, fm = np.zeros(5), np.zeros(5)
    L.orc_op_counts_reset()
    L.orc_lf_split(q.ctypes.data_as(DP), 5, k.ctypes.data_as(DP), 2.0, fp.ctypes.data_as(DP),
                   fm.ctypes.data_as(DP))
    np.testing.assert_array_equal(counts(L), LF5)


def _box(nx, ny, nz):
    c = cases.mixing_layer_case([nx - 1], [ny - 1], [nz - 1])
    return c


@pytest.mark.parametrize("dims", [(7, 9, 6), (12, 5, 8)])
def test_operator_counts_follow_the_pencil_structure(dims):
    case = _box(*dims)
    o = O.Oracle(case, tangent="count")
    o.setup_mixing_layer()
    o.initialize()
    L, g = o.L, o.ngg
    m = np.array([case["blocks"][0][k] for k in ("mx", "my", "mz")])
    inv = np.zeros(5, dtype=np.int64)
    vis = np.zeros(5, dtype=np.int64)
    for d in range(3):
        pencils = int(np.prod(m)) // m[d]
        inv += pencils * ((m[d] + 2 * g) * INV_POINT + (m[d] + 1) * INV_FACE + m[d] * CELL)
        vis += pencils * ((m[d] + 1) * VISC_FACE + m[d] * CELL)
    L.orc_zero_dq(o.ctx, 0)
    L.orc_op_counts_reset()
    L.orc_inviscid(o.ctx, 0)
    np.testing.assert_array_equal(counts(L), inv)
    L.orc_op_counts_reset()
    L.orc_viscous(o.ctx, 0)
    np.testing.assert_array_equal(counts(L), vis)
