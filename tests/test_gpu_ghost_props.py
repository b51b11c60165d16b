"""Properties of the ghost layers after a batched stage computed by the ghost
ops themselves (GHOST_PROPS_IN_OPS, the default) against one more pass over
every ghost layer (GHOST_PROPS_PASS) on the GPU (needs a GPU): the states
after n batched RK3 steps -- cv, primitives, viscosity and Mach number on the
padded blocks -- are bit-identical, for periodic, outflow, farfield and
inflow (white-noise) faces, rotated inner faces, the sponge layer, and the
24-block case.  Against the oracle the default is held to the step tests'
1e-9 here too.
"""
import numpy as np
import pytest

from _util import O, cases, oracle_for, oracle_pair, pkg, rel_err
from test_gpu_connectivity import _global_bv, _solver_pair

pytestmark = pytest.mark.gpu


def _cases():
    out = {
        "periodic-outflow-12": cases.small_multiblock_case(),
        "farfield-12": cases.small_multiblock_case(y_bc=cases.BC_FARFIELD),
        "inflow-noise-8": cases.mixing_layer_case([9, 12], [8, 7], [6, 5], spatial=True, inflow_noise=1e-2),
        "rotated-pair": cases.rotated_pair(cases.mixing_layer_case([23], [15], [13]), 10, split_k=7),
        "scheme-72-12": cases.small_multiblock_case(scheme=72),
    }
    sp = cases.mixing_layer_case([19], [15], [11], spatial=True, inflow_noise=1e-2)
    sp["sponge"] = cases.sponge_layer(sp, sides=(1, 2, 3), depth=0.3, coefficient=2.0, jacobian_scaled=1)
    out["sponge"] = sp
    return out


CASES = _cases()
FIELDS = (pkg.FIELD_CV, pkg.FIELD_BV, pkg.FIELD_MUL, pkg.FIELD_MACH)


def _single_bv():
    """The rotated pair's state: that of the single block it was cut from."""
    return _global_bv(cases.mixing_layer_case([23], [15], [13]))


def _solver(case, name=None):
    if name == "rotated-pair":
        return _solver_pair(case, _single_bv())
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    s.initialize()
    return s


def _run(case, name, mode, n):
    s = _solver(case, name)
    s.set_ghost_props(mode)
    dt = s.rk3(n, pkg.MODE_BATCHED)
    states = [[s.get(b, f) for f in FIELDS] for b in range(len(case["blocks"]))]
    s.close()
    return dt, states


@pytest.mark.parametrize("name", list(CASES))
def test_props_in_ops_equal_props_pass(name):
    case = CASES[name]
    nb = len(case["blocks"])
    n = 6
    print(f"[gpu] ghost properties in the ops vs a pass, {name}, {nb} blocks, {n} batched steps", flush=True)
    dt_a, st_a = _run(case, name, pkg.GHOST_PROPS_PASS, n)
    dt_b, st_b = _run(case, name, pkg.GHOST_PROPS_IN_OPS, n)
    assert dt_a == dt_b
    for b in range(nb):
        for f, (x, y) in zip(FIELDS, zip(st_a[b], st_b[b])):
            assert np.array_equal(x, y), (name, b, f, np.max(np.abs(x - y)))
    if name in ("inflow-noise-8", "rotated-pair"):
        o = oracle_pair(case, _single_bv()) if name == "rotated-pair" else oracle_for(case)
        o.step(n)
        for b in range(nb):
            e = rel_err(st_b[b][1], o.array(b, O.BV))
            print(f"  block {b} bv rel err vs oracle {e}", flush=True)
            assert np.all(e <= 1e-9), (b, e)


def test_props_in_ops_equal_props_pass_24_blocks():
    case = cases.batched_case()
    nb = len(case["blocks"])
    print(f"[gpu] ghost properties in the ops vs a pass, 24-block case ({case['n_cells']} cells), 2 steps",
          flush=True)
    dt_a, st_a = _run(case, None, pkg.GHOST_PROPS_PASS, 2)
    dt_b, st_b = _run(case, None, pkg.GHOST_PROPS_IN_OPS, 2)
    assert dt_a == dt_b
    for b in range(nb):
        for f, (x, y) in zip(FIELDS, zip(st_a[b], st_b[b])):
            assert np.array_equal(x, y), (b, f)


def test_mode_is_checked():
    s = _solver(CASES["periodic-outflow-12"])
    with pytest.raises(pkg.OstcError, match="unknown ghost properties mode"):
        s.set_ghost_props(5)
    s.close()
