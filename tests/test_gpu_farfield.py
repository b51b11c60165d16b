"""Farfield boundary (label 4) of the HIP library vs the CPU oracle (needs a
GPU).

* FarField constructor (host code of the library) = the oracle's, bitwise,
  for the three input combinations;
* one boundary application on a box whose normal velocity sweeps through
  supersonic/subsonic inflow and outflow along x, on a tilted face metric:
  bv and cv of the whole padded block <= 1e-13 (pow and sqrt are the only
  library routines that differ between device and host);
* 10 RK3 steps of the mixing layer with farfield y faces, single block and
  3x2x2 blocks, per-block and batched paths: primitives <= 1e-9.
"""
import numpy as np
import pytest

from _util import O, cases, interior, oracle_for, pkg, rel_err
from test_oracle_farfield import HI, LO, _case
from test_oracle_rhs import exact_grid, set_state

pytestmark = pytest.mark.gpu

SPECS = [dict(mach=0.7, pressure=9.0e4, temperature=280.0, u=0.6, v=0.8, w=0.0),
         dict(density=1.1, velocity=250.0, pressure=1.0e5),
         dict(mach=2.0, temperature=220.0, reynolds_number=3.0e6)]


@pytest.mark.parametrize("spec", range(len(SPECS)))
def test_farfield_state_matches_oracle(spec):
    a = pkg.farfield_state_air(**SPECS[spec])
    b = O.farfield_state_air(**SPECS[spec])
    for name, _ in pkg.Farfield._fields_[:10]:
        assert getattr(a, name) == getattr(b, name), name


def test_farfield_application_matches_oracle():
    print("[gpu] farfield BC on a swept box", flush=True)
    n = 11
    case = _case(n, LO, HI)
    o = O.Oracle(case)
    exact_grid(o)
    met = o.array(0, O.METRIC)
    met[3][...] = 0.3
    met[5][...] = -0.5
    rho, p = 1.15, 1.0e5
    c = np.sqrt(1.4 * p / rho)
    shp = o.shape(0, O.BV)
    v = np.broadcast_to(np.linspace(-2.2 * c, 2.2 * c, shp[2]), shp).copy()
    set_state(o, rho, 120.0, v, -40.0, p)
    s = pkg.Solver(case)
    s.set(0, pkg.FIELD_JAC, o.array(0, O.JAC))
    s.set(0, pkg.FIELD_METRIC, o.array(0, O.METRIC))
    s.set(0, pkg.FIELD_BV, o.array(0, O.BV))
    s.set(0, pkg.FIELD_CV, o.array(0, O.CV))
    assert O.lib().orc_apply_bc(o.ctx, 0) == 0
    s.apply_boundary_conditions(0)
    for f, w in ((pkg.FIELD_BV, O.BV), (pkg.FIELD_CV, O.CV)):
        e = rel_err(s.get(0, f), o.array(0, w))
        print(f"  field {f} rel err {e}", flush=True)
        assert np.all(e <= 1e-13), (f, e)
    s.close()


@pytest.mark.parametrize("kind", ["single", "multi"])
@pytest.mark.parametrize("mode", [pkg.MODE_PLAIN, pkg.MODE_BATCHED], ids=["plain", "batched"])
def test_mixing_layer_with_farfield_faces(kind, mode):
    n = 10
    case = (cases.single_block_case(16, y_bc=cases.BC_FARFIELD) if kind == "single"
            else cases.small_multiblock_case(y_bc=cases.BC_FARFIELD))
    print(f"[gpu] farfield mixing layer, {kind}, mode {mode}, {n} steps", flush=True)
    o = oracle_for(case)
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    s.initialize()
    dt_o = o.step(n)
    dt_g = s.rk3(n, mode)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (dt_g, dt_o)
    g = s.ngg
    for b in range(len(case["blocks"])):
        e = rel_err(interior(s.get(b, pkg.FIELD_BV), g), interior(o.array(b, O.BV), g))
        print(f"  block {b} bv rel err {e}", flush=True)
        assert np.all(e <= 1e-9), (b, e)
    s.close()
