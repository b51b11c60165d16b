"""The spatially developing mixing layer on the GPU (needs a GPU): inflow
faces (label 5, the mixing-layer profile with white noise) at x-, outflow at
x+, against the oracle on the same seeded inputs.

Tolerance (per variable, max|gpu - oracle| / max|oracle|): primitives and
mass fractions after n RK3 steps <= 1e-9.  The noise deviates come from the
same counter-based generator on both sides (device tanh/log/sin/cos against
glibc: ulp-level differences, inside the tolerance).
"""
import numpy as np
import pytest

from _util import O, cases, interior, oracle_for, pkg, rel_err

pytestmark = pytest.mark.gpu

BV_TOL = 1e-9


@pytest.mark.parametrize("blocks,mode", [((23, 15, 11), "plain"), ((23, 15, 11), "batched"),
                                         (([9, 12], [8, 7], [6, 5]), "batched")])
def test_air_spatial_layer_steps(blocks, mode):
    n = 8
    if isinstance(blocks[0], int):
        case = cases.mixing_layer_case([blocks[0]], [blocks[1]], [blocks[2]], spatial=True, inflow_noise=1e-2)
    else:
        case = cases.mixing_layer_case(*blocks, spatial=True, inflow_noise=1e-2)
    print(f"[gpu] air spatial layer, {len(case['blocks'])} blocks, {mode}, {n} steps", flush=True)
    o = oracle_for(case)
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    s.initialize()
    dt_o = o.step(n)
    dt_g = s.rk3(n, pkg.MODE_PLAIN if mode == "plain" else pkg.MODE_BATCHED)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (dt_g, dt_o)
    g = s.ngg
    for b in range(len(case["blocks"])):
        e = rel_err(s.get(b, pkg.FIELD_BV), o.array(b, O.BV))
        print(f"  block {b} bv rel err {e}", flush=True)
        assert np.all(e <= BV_TOL), (b, e)
    # the inflow face carries noise: its u differs from the clean profile
    bv = s.get(0, pkg.FIELD_BV)
    assert np.std(interior(bv, g)[1][:, :, 0]) > 0
    s.close()


@pytest.mark.parametrize("blocks", [(11, 17, 9), ([6, 6], [8, 8], [4, 5])])
def test_mixture_spatial_layer_steps(blocks):
    import mech
    import oracle_fr as OF
    n = 4
    tg = pkg.mixture.load_mechanism(*pkg.mixture.H2_AIR_FILES)
    to = mech.load(*pkg.mixture.H2_AIR_FILES)
    if isinstance(blocks[0], int):
        case = cases.reacting_case([blocks[0]], [blocks[1]], [blocks[2]], to, spatial=True, inflow_noise=1e-2)
    else:
        case = cases.reacting_case(*blocks, to, spatial=True, inflow_noise=1e-2)
    print(f"[gpu] mixture spatial layer, {len(case['blocks'])} blocks, {n} steps", flush=True)
    o = OF.FrOracle(case, to, n_threads=8)
    o.setup_mixing_layer()
    o.initialize()
    s = pkg.mixture.FrSolver(case, tg)
    s.setup_mixing_layer()
    s.initialize()
    dt_o = o.step(n)
    dt_g = s.rk3(n)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (dt_g, dt_o)
    for b in range(len(case["blocks"])):
        for f, w in ((pkg.mixture.FIELD_BV, OF.BV), (pkg.mixture.FIELD_SV, OF.SV)):
            e = rel_err(s.get(b, f), o.array(b, w))
            print(f"  block {b} field {f} rel err {e}", flush=True)
            assert np.all(e <= BV_TOL), (b, f, e)
    s.close()


def test_inflow_records_are_checked():
    print("[gpu] inflow record checks", flush=True)
    case = cases.mixing_layer_case([11], [9], [7], spatial=True)
    bad = dict(case)
    bad["boundaries"] = [dict(f) for f in case["boundaries"]]
    for f in bad["boundaries"]:
        if f["type"] == cases.BC_INFLOW:
            f["inflow"] = dict(f["inflow"], inflow_type=1)   # the profile inflow needs profile files
    with pytest.raises(pkg.OstcError, match="inflow_type"):
        pkg.Solver(bad)
    for f in bad["boundaries"]:
        if f["type"] == cases.BC_INFLOW:
            f["inflow"] = dict(f["inflow"], inflow_type=2, delta_omega=0.0)
    with pytest.raises(pkg.OstcError, match="delta_omega"):
        pkg.Solver(bad)
