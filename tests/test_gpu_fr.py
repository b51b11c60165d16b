"""Finite-rate mixture path on the GPU against the CPU oracle (needs a GPU).

H2/air (9 species, 19 reactions, the reference's H2_mech.inp with its
therm.dat / tran.dat), temporal mixing layer with a species and temperature
step at y = 0:

* properties over the padded domain (cp, gamma, c, mu, lambda, rho D_k) and
  cv after initialisation: <= 1e-13 relative;
* the stage residual dq for component-wise (51) and characteristic (52)
  WENO-Z5, inviscid / viscous, with and without chemistry and the
  positivity limiter: <= 1e-12 relative per variable for the flow rows,
  <= 1e-10 for the species rows when the chemical source is on;
* whole SSP-RK3 steps (direct launches and the captured graph), one block and
  a multi-block partition with inner faces: primitives and mass fractions
  <= 1e-9 after the steps.
The GPU runs on tables from the library's reader, the oracle on tables from
its own reader (the two agree bitwise, tests/test_oracle_fr.py).
"""
import numpy as np
import pytest

from _util import pkg, rel_err
import mech
import oracle_fr as OF

pytestmark = pytest.mark.gpu

MX = pkg.mixture
FILES = MX.H2_AIR_FILES


@pytest.fixture(scope="module")
def tables():
    return MX.load_mechanism(*FILES), mech.load(*FILES)


def _pair(case, tables):
    tg, to = tables
    o = OF.FrOracle(case, to)
    o.setup_mixing_layer()
    o.initialize()
    s = MX.FrSolver(case, tg)
    s.setup_mixing_layer()
    s.initialize()
    return o, s


def _case(tables, nx=12, ny=16, nz=10, **kw):
    return pkg.cases.reacting_single_block(nx, ny, nz, tables[0], **kw)


def test_initial_state_and_properties(tables):
    print("[gpu] fr initial state and properties", flush=True)
    o, s = _pair(_case(tables), tables)
    for name, fo, fg, tol in (("cv", OF.CV, MX.FIELD_CV, 1e-13), ("bv", OF.BV, MX.FIELD_BV, 0),
                              ("sv", OF.SV, MX.FIELD_SV, 0), ("mu", OF.MUL, MX.FIELD_MUL, 1e-13),
                              ("lambda", OF.LAMBDA, MX.FIELD_LAMBDA, 1e-13), ("cp", OF.CP, MX.FIELD_CP, 1e-14),
                              ("gamma", OF.GAMMA, MX.FIELD_GAMMA, 1e-14), ("c", OF.C_SOUND, MX.FIELD_C, 1e-14),
                              ("rhoD", OF.RHOD, MX.FIELD_RHOD, 1e-13), ("mach", OF.MACH, MX.FIELD_MACH, 1e-13)):
        err = rel_err(s.get(0, fg), o.array(0, fo))
        print(f"  {name}: {err.max():.2e}", flush=True)
        assert np.all(err <= tol), (name, err)
    s.close()


@pytest.mark.parametrize("scheme", [51, 52])
@pytest.mark.parametrize("viscous", [0, 2])
@pytest.mark.parametrize("reaction", [0, 1])
@pytest.mark.parametrize("pp", [0, 1])
def test_stage_residual(tables, scheme, viscous, reaction, pp):
    print(f"[gpu] fr dq scheme {scheme} viscous {viscous} reaction {reaction} pp {pp}", flush=True)
    case = _case(tables, scheme=scheme, viscous_order=viscous, reaction=reaction, positive_preserving=pp)
    o, s = _pair(case, tables)
    dt = 2e-9
    dq_o = o.rhs(0, dt).copy()
    s.compute_rhs(0, dt)
    dq_g = s.get(0, MX.FIELD_DQ)
    err = rel_err(dq_g, dq_o)
    print("  per-variable:", " ".join(f"{e:.1e}" for e in err), flush=True)
    assert np.all(err[:5] <= 1e-12)
    assert np.all(err[5:] <= (1e-10 if reaction else 1e-12))
    s.close()


def _compare_state(o, s, nblocks, tol, label):
    g = 3
    worst = 0.0
    for b in range(nblocks):
        for fo, fg in ((OF.BV, MX.FIELD_BV), (OF.SV, MX.FIELD_SV)):
            a = s.get(b, fg)[:, g:-g, g:-g, g:-g]
            r = o.array(b, fo)[:, g:-g, g:-g, g:-g]
            err = rel_err(a, r)
            worst = max(worst, float(err.max()))
            assert np.all(err <= tol), (label, b, fo, err)
    print(f"  {label}: worst relative error {worst:.2e}", flush=True)


@pytest.mark.parametrize("mode", [MX.MODE_DIRECT, MX.MODE_GRAPH], ids=["direct", "graph"])
@pytest.mark.parametrize("scheme", [51, 52])
def test_steps_single_block(tables, mode, scheme):
    print(f"[gpu] fr 10 steps single block scheme {scheme} mode {mode}", flush=True)
    case = _case(tables, scheme=scheme, positive_preserving=1)
    o, s = _pair(case, tables)
    dt_o = o.step(10)
    dt_g = s.rk3(10, mode)
    print(f"  dt oracle {dt_o:.15e} gpu {dt_g:.15e}", flush=True)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o
    _compare_state(o, s, 1, 1e-9, "after 10 steps")
    s.close()


def test_steps_multiblock(tables):
    print("[gpu] fr 5 steps on a 2 x 2 x 2 partition", flush=True)
    case = pkg.cases.reacting_case([7, 6], [8, 9], [5, 6], tables[0], scheme=52)
    o, s = _pair(case, tables)
    o.step(5)
    s.rk3(5, MX.MODE_GRAPH)
    _compare_state(o, s, len(case["blocks"]), 1e-9, "after 5 steps")
    s.close()


def test_fixed_dt_and_stage_pieces(tables):
    """ostc_fr_stage_* with the caller's dt, then whole steps with the fixed
    dt of the setup: the pieces must not leave their dt behind."""
    print("[gpu] fr fixed dt after stage pieces", flush=True)
    case = _case(tables, scheme=52)
    case["fixed_time_step"] = 1
    case["dt"] = 3e-9
    o, s = _pair(case, tables)
    s.compute_rhs(0, 1e-7)
    s.update_cv_and_bv_rk(0, 0, 1e-7)
    s.setup_mixing_layer()
    s.initialize()
    o.step(3)
    s.rk3(3, MX.MODE_GRAPH)
    _compare_state(o, s, 1, 1e-9, "fixed dt, 3 steps")
    s.close()
