"""HIP library vs the CPU oracle on the same seeded inputs (needs a GPU).

Tolerances (per variable, max|gpu - oracle| / max|oracle|):
  dq (inviscid + viscous)            <= 1e-12
  primitives after n RK3 steps       <= 1e-9
  state after initialisation         <= 1e-13
Both the HIP build and the oracle round every product and sum separately in
the written order (-ffp-contract=off), so dq agrees bit for bit in practice;
the tolerances above are the gates.
"""
import numpy as np
import pytest

from _util import O, cases, interior, oracle_for, pkg, rel_err

pytestmark = pytest.mark.gpu

DQ_TOL = 1e-12
BV_TOL = 1e-9


def _solver(case):
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    return s


def _load_oracle_state(s, o, nblocks):
    for b in range(nblocks):
        s.set(b, pkg.FIELD_CV, o.array(b, O.CV))
        s.set(b, pkg.FIELD_BV, o.array(b, O.BV))
        s.set(b, pkg.FIELD_MUL, o.array(b, O.MUL))


SINGLE = cases.single_block_case(20)
MULTI = cases.small_multiblock_case()


@pytest.mark.parametrize("case", [SINGLE, MULTI], ids=["single", "multi"])
def test_setup_and_initialize(case):
    print(f"[gpu] setup/initialize, {len(case['blocks'])} blocks", flush=True)
    o = oracle_for(case)
    s = _solver(case)
    g = s.ngg
    for b in range(len(case["blocks"])):
        np.testing.assert_array_equal(s.get(b, pkg.FIELD_JAC)[0], o.array(b, O.JAC)[0])
        np.testing.assert_array_equal(s.get(b, pkg.FIELD_METRIC), o.array(b, O.METRIC))
    s.initialize()
    for b in range(len(case["blocks"])):
        for f, w, tol in ((pkg.FIELD_BV, O.BV, 1e-13), (pkg.FIELD_CV, O.CV, 1e-13), (pkg.FIELD_MUL, O.MUL, 1e-13)):
            e = rel_err(s.get(b, f), o.array(b, w))
            assert np.all(e <= tol), (b, f, e)
    s.close()


@pytest.mark.parametrize("case", [SINGLE, MULTI], ids=["single", "multi"])
@pytest.mark.parametrize("path", ["plain", "batched"])
def test_rhs_dq(case, path):
    print(f"[gpu] rhs dq {path}, {len(case['blocks'])} blocks", flush=True)
    o = oracle_for(case)
    s = _solver(case)
    _load_oracle_state(s, o, len(case["blocks"]))
    for b in range(len(case["blocks"])):
        o.rhs(b)
        if path == "plain":
            s.compute_rhs(b)
        else:
            s.compute_rhs_batched(b)
        dq_g = s.get(b, pkg.FIELD_DQ)
        dq_o = o.array(b, O.DQ)
        e = rel_err(dq_g, dq_o)
        print(f"  block {b} dq rel err {e}", flush=True)
        assert np.all(e <= DQ_TOL), (b, e)
    s.close()


@pytest.mark.parametrize("index", range(6))
def test_rhs_dq_every_tile_shape(index):
    case = MULTI
    print(f"[gpu] batched rhs, tile table entry {index}", flush=True)
    o = oracle_for(case)
    s = _solver(case)
    _load_oracle_state(s, o, len(case["blocks"]))
    s.set_tile(index, 0)
    for b in range(len(case["blocks"])):
        o.rhs(b)
        s.compute_rhs_batched(b)
        e = rel_err(s.get(b, pkg.FIELD_DQ), o.array(b, O.DQ))
        assert np.all(e <= DQ_TOL), (index, b, e)
    s.close()


@pytest.mark.parametrize("case", [SINGLE, MULTI], ids=["single", "multi"])
@pytest.mark.parametrize("mode", [pkg.MODE_PLAIN, pkg.MODE_BATCHED], ids=["plain", "batched"])
def test_steps(case, mode):
    n = 10
    print(f"[gpu] {n} RK3 steps mode {mode}, {len(case['blocks'])} blocks", flush=True)
    o = oracle_for(case)
    s = _solver(case)
    s.initialize()
    dt_o = o.step(n)
    dt_g = s.rk3(n, mode)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (dt_g, dt_o)
    g = s.ngg
    for b in range(len(case["blocks"])):
        e = rel_err(interior(s.get(b, pkg.FIELD_BV), g), interior(o.array(b, O.BV), g))
        print(f"  block {b} bv rel err {e}", flush=True)
        assert np.all(e <= BV_TOL), (b, e)
    s.close()


def test_persistent_grid_matches_one_wg_per_tile():
    case = MULTI
    print("[gpu] persistent grid vs one workgroup per tile", flush=True)
    s1, s2 = _solver(case), _solver(case)
    s1.initialize()
    s2.initialize()
    s2.set_tile(0, 2)
    s1.rk3(3, pkg.MODE_BATCHED)
    s2.rk3(3, pkg.MODE_BATCHED)
    for b in range(len(case["blocks"])):
        np.testing.assert_array_equal(s1.get(b, pkg.FIELD_BV), s2.get(b, pkg.FIELD_BV))
    s1.close()
    s2.close()


@pytest.mark.parametrize("split", [0, 1], ids=["one-kernel", "split-stage"])
def test_batched_launch_configs_match_oracle(split):
    case = MULTI
    n = 5
    print(f"[gpu] batched steps, split={split}, persistent grid", flush=True)
    o = oracle_for(case)
    s = _solver(case)
    s.initialize()
    s.set_tile(1, 2, split)
    dt_o = o.step(n)
    dt_g = s.rk3(n, pkg.MODE_BATCHED)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o
    g = s.ngg
    for b in range(len(case["blocks"])):
        e = rel_err(interior(s.get(b, pkg.FIELD_BV), g), interior(o.array(b, O.BV), g))
        assert np.all(e <= BV_TOL), (b, e)
    s.close()


_C0_CACHE = {}


def _config0_oracle(scheme):
    """Oracle state after 100 SSP-RK3 steps of the 64^3 single-block mixing
    layer (periodic x/z, outflow y, CFL 0.5)."""
    if scheme not in _C0_CACHE:
        case = cases.single_block_case(64, scheme=scheme)
        o = oracle_for(case)
        dt = o.step(100)
        _C0_CACHE[scheme] = (case, dt, o.array(0, O.BV).copy())
    return _C0_CACHE[scheme]


@pytest.mark.parametrize("scheme", [51, 50], ids=["wenoz5", "weno5js"])
@pytest.mark.parametrize("mode", [pkg.MODE_PLAIN, pkg.MODE_BATCHED], ids=["plain", "batched"])
def test_single_block_64_cubed_100_steps(scheme, mode):
    print(f"[gpu] 64^3, 100 steps, scheme {scheme}, mode {mode}", flush=True)
    case, dt_o, bv_o = _config0_oracle(scheme)
    s = _solver(case)
    s.initialize()
    dt_g = s.rk3(100, mode)
    g = s.ngg
    e = rel_err(interior(s.get(0, pkg.FIELD_BV), g), interior(bv_o, g))
    print(f"  dt gpu {dt_g:.15e} oracle {dt_o:.15e}; primitive rel err {e}", flush=True)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o
    assert np.all(e <= BV_TOL), e
    s.close()


def test_residual_matches_oracle():
    """ostc_store_last_step / ostc_residual (src/TimeAdvanceFunc.cu:6-37,
    src/Residual.cuh:12-187) on identical states: only the summation order
    differs."""
    print("[gpu] residual of two oracle states", flush=True)
    case = MULTI
    nb = len(case["blocks"])
    o = oracle_for(case)
    s = _solver(case)
    _load_oracle_state(s, o, nb)
    s.store_last_step()
    assert O.lib().orc_store_last_step(o.ctx) == 0
    o.step(2)
    _load_oracle_state(s, o, nb)
    got, want = s.compute_residual(), o.residual()
    print(f"  residual gpu {got} oracle {want}", flush=True)
    np.testing.assert_allclose(got, want, rtol=1e-12)
    assert np.all(want > 0)
    s.close()
