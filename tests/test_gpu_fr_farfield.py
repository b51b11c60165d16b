"""Farfield boundary (label 4) of the finite-rate mixture path vs the CPU
oracle (needs a GPU).

* FarField constructor with species (host code of the library) = the
  oracle's, bitwise, for the three input combinations;
* initialisation (properties, boundaries, properties) of a box whose normal
  velocity sweeps supersonic/subsonic inflow and outflow on a tilted face,
  interior and far-field compositions differing: bv, mass fractions, cv and
  gamma of the padded block <= 1e-12;
* SSP-RK3 steps of the H2/air mixing layer with farfield y faces (fuel stream
  above, oxidizer below), one block (direct and graph) and a 2x2x2
  partition: primitives and mass fractions <= 1e-9.
"""
import numpy as np
import pytest

from _util import pkg, rel_err
import mech
import oracle_fr as OF
from test_gpu_fr import _compare_state, _pair
from test_oracle_fr_farfield import HI, LO, Y_AIR, Y_FUEL, Y_IN, _case, _gamma, R_U

pytestmark = pytest.mark.gpu

MX = pkg.mixture


@pytest.fixture(scope="module")
def tables():
    return MX.load_mechanism(*MX.H2_AIR_FILES), mech.load(*MX.H2_AIR_FILES)


SPECS = [(dict(mach=0.8, pressure=9.0e4, temperature=1500.0, u=0.6, v=0.8), Y_AIR),
         (dict(density=0.4, velocity=300.0, pressure=1.0e5), Y_FUEL),
         (dict(mach=1.5, temperature=400.0, reynolds_number=2.0e6), Y_FUEL)]


@pytest.mark.parametrize("spec", range(len(SPECS)))
def test_mixture_farfield_state_matches_oracle(tables, spec):
    kw, y = SPECS[spec]
    a = MX.farfield_state(tables[0], y, **kw)
    b = OF.farfield_state(tables[1], y, **kw)
    for name, _ in pkg.Farfield._fields_[:10]:
        assert getattr(a, name) == getattr(b, name), name
    assert list(a.sv) == list(b.sv)


def test_mixture_farfield_initialisation_matches_oracle(tables):
    print("[gpu] fr farfield initialisation on a swept box", flush=True)
    tg, to = tables
    case = _case(to, 9, LO, HI)
    y_in = Y_IN.copy()
    y_in[8] = 1 - y_in[:8].sum()
    o = OF.FrOracle(case, to, n_threads=4)
    jac, met = o.array(0, OF.JAC), o.array(0, OF.METRIC)
    jac[...] = 0.125
    met[...] = 0.0
    met[0], met[4], met[8] = 2.0, 2.0, 2.0
    met[3], met[5] = 0.3, -0.5
    bv, sv = o.array(0, OF.BV), o.array(0, OF.SV)
    sv[...] = y_in[:, None, None, None]
    g0, mw0 = _gamma(to, 900.0, y_in)
    rho, p = 0.6, 0.6 * R_U / mw0 * 900.0
    c = np.sqrt(g0 * p / rho)
    shp = bv.shape[1:]
    bv[0], bv[1], bv[3], bv[4], bv[5] = rho, 120.0, -40.0, p, 900.0
    bv[2] = np.broadcast_to(np.linspace(-2.2 * c, 2.2 * c, shp[2]), shp)
    s = MX.FrSolver(case, tg)
    s.set_geometry(0, o.array(0, OF.JAC), o.array(0, OF.METRIC))
    s.set(0, MX.FIELD_BV, bv)
    s.set(0, MX.FIELD_SV, sv)
    o.initialize()
    s.initialize()
    for name, fo, fg in (("bv", OF.BV, MX.FIELD_BV), ("sv", OF.SV, MX.FIELD_SV), ("cv", OF.CV, MX.FIELD_CV),
                         ("gamma", OF.GAMMA, MX.FIELD_GAMMA)):
        e = rel_err(s.get(0, fg), o.array(0, fo))
        print(f"  {name} rel err {e.max():.2e}", flush=True)
        assert np.all(e <= 1e-12), (name, e)
    s.close()


@pytest.mark.parametrize("mode", [MX.MODE_DIRECT, MX.MODE_GRAPH], ids=["direct", "graph"])
def test_mixing_layer_with_farfield_faces(tables, mode):
    print(f"[gpu] fr farfield mixing layer, 10 steps, mode {mode}", flush=True)
    case = pkg.cases.reacting_single_block(12, 16, 10, tables[0], scheme=52, positive_preserving=1,
                                           y_bc=pkg.cases.BC_FARFIELD)
    o, s = _pair(case, tables)
    dt_o = o.step(10)
    dt_g = s.rk3(10, mode)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (dt_g, dt_o)
    _compare_state(o, s, 1, 1e-9, "farfield, 10 steps")
    s.close()


def test_mixing_layer_with_farfield_faces_multiblock(tables):
    print("[gpu] fr farfield mixing layer, 2x2x2 blocks, 5 steps", flush=True)
    case = pkg.cases.reacting_case([7, 6], [8, 9], [5, 6], tables[0], scheme=52, y_bc=pkg.cases.BC_FARFIELD)
    o, s = _pair(case, tables)
    o.step(5)
    s.rk3(5, MX.MODE_GRAPH)
    _compare_state(o, s, len(case["blocks"]), 1e-9, "farfield, 5 steps")
    s.close()
