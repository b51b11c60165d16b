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
    g = s.ngg
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


def _mixed_state(case, tables, seed, radicals=1.0, t_mean=1400.0):
    """A smooth random state with every species present at t_mean +- 20 %,
    on the padded domain of block 0.  radicals scales the mass fractions of
    H, H2O2, HO2, O and OH (1: as much as the stable species -- a state that
    explicit chemistry cannot step; 1e-5: traces that it can)."""
    rng = np.random.default_rng(seed)
    b = case["blocks"][0]
    g = 4 if case["inviscid_scheme"] in (71, 72) else 3
    shp = (b["mz"] + 2 * g, b["my"] + 2 * g, b["mx"] + 2 * g)
    k, j, i = np.meshgrid(*(np.arange(n) for n in shp), indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, (16, 3))
    w = lambda q: np.sin(0.7 * i + ph[q, 0]) * np.cos(0.5 * j + ph[q, 1]) + 0.5 * np.sin(0.6 * k + ph[q, 2])
    y = np.array([1.0 + 0.6 * w(q) for q in range(9)])
    for q in (0, 3, 4, 5, 7):  # H, H2O2, HO2, O, OH
        y[q] *= radicals
    sv = y / y.sum(axis=0)
    bv = np.empty((6, *shp))
    bv[0] = 0.3 * (1 + 0.2 * w(9))
    bv[1], bv[2], bv[3] = 200 * w(10), 90 * w(11), 70 * w(12)
    bv[5] = t_mean * (1 + 0.2 * w(13))
    mw = np.array(tables[1].mw[:9])
    bv[4] = bv[0] * 8314.462618 * np.tensordot(1 / mw, sv, axes=1) * bv[5]
    return bv, sv


def _pair_state(case, tables, bv, sv):
    tg, to = tables
    o = OF.FrOracle(case, to)
    o.setup_mixing_layer()
    o.array(0, OF.BV)[...] = bv
    o.array(0, OF.SV)[...] = sv
    o.initialize()
    s = MX.FrSolver(case, tg)
    s.setup_mixing_layer()
    s.set(0, MX.FIELD_BV, bv)
    s.set(0, MX.FIELD_SV, sv)
    s.initialize()
    return o, s


@pytest.mark.parametrize("scheme", [51, 52])
@pytest.mark.parametrize("pp", [0, 1])
def test_stage_residual_reacting_mixture(tables, scheme, pp):
    print(f"[gpu] fr dq on a premixed radical state, scheme {scheme} pp {pp}", flush=True)
    case = _case(tables, scheme=scheme, positive_preserving=pp)
    bv, sv = _mixed_state(case, tables, 11 + scheme + pp)
    o, s = _pair_state(case, tables, bv, sv)
    for name, fo, fg in (("mu", OF.MUL, MX.FIELD_MUL), ("lambda", OF.LAMBDA, MX.FIELD_LAMBDA),
                         ("rhoD", OF.RHOD, MX.FIELD_RHOD), ("c", OF.C_SOUND, MX.FIELD_C)):
        err = rel_err(s.get(0, fg), o.array(0, fo))
        print(f"  {name}: {err.max():.2e}", flush=True)
        assert np.all(err <= 1e-13), name
    dt = 4e-9
    dq_o = o.rhs(0, dt).copy()
    s.compute_rhs(0, dt)
    err = rel_err(s.get(0, MX.FIELD_DQ), dq_o)
    print("  dq per-variable:", " ".join(f"{e:.1e}" for e in err), flush=True)
    chem = np.abs(dq_o[5:]).max()
    assert chem > 0
    assert np.all(err[:5] <= 1e-12)
    assert np.all(err[5:] <= 1e-10)
    # the chemical source alone, species rows
    o.L.orf_zero_dq(o.ctx, 0)
    o.L.orf_source(o.ctx, 0)
    src = o.array(0, OF.DQ)[5:].copy()
    assert np.abs(src).max() > 1e-3 * chem  # the source is a real part of dq here
    s.close()


def test_steps_reacting_mixture(tables):
    print("[gpu] fr 5 steps from a premixed radical state", flush=True)
    case = _case(tables, scheme=52, positive_preserving=1)
    bv, sv = _mixed_state(case, tables, 21, radicals=1e-5, t_mean=1000.0)
    o, s = _pair_state(case, tables, bv, sv)
    o.step(5)
    s.rk3(5, MX.MODE_GRAPH)
    _compare_state(o, s, 1, 1e-9, "reacting, 5 steps")
    s.close()


@pytest.mark.parametrize("scheme", [2, 3, 4], ids=["roe", "ausm+", "hllc"])
@pytest.mark.parametrize("recon", [1, 2, 3], ids=["1st", "muscl", "nnd2"])
def test_stage_residual_primitive_reconstruction(tables, scheme, recon):
    print(f"[gpu] fr dq, primitive reconstruction {recon}, Riemann solver {scheme}", flush=True)
    case = _case(tables, scheme=scheme)
    case["reconstruction"] = recon
    for label, state in (("layer", None), ("premixed", _mixed_state(case, tables, 31 + recon))):
        o, s = _pair(case, tables) if state is None else _pair_state(case, tables, *state)
        dq_o = o.rhs(0, 2e-9).copy()
        s.compute_rhs(0, 2e-9)
        err = rel_err(s.get(0, MX.FIELD_DQ), dq_o)
        print(f"  {label} dq per-variable:", " ".join(f"{e:.1e}" for e in err), flush=True)
        assert np.all(err[:5] <= 1e-12)
        assert np.all(err[5:] <= 1e-10)
        s.close()


@pytest.mark.parametrize("scheme", [2, 3, 4], ids=["roe", "ausm+", "hllc"])
def test_steps_primitive_reconstruction(tables, scheme):
    print(f"[gpu] fr 10 steps, MUSCL + Riemann solver {scheme}", flush=True)
    case = _case(tables, scheme=scheme)
    case["reconstruction"] = 2
    o, s = _pair(case, tables)
    o.step(10)
    s.rk3(10, MX.MODE_GRAPH)
    _compare_state(o, s, 1, 1e-9, "10 steps")
    s.close()


@pytest.mark.parametrize("scheme", [71, 72], ids=["wenoz7-cp", "wenoz7-char"])
@pytest.mark.parametrize("reaction", [0, 1])
def test_stage_residual_weno7(tables, scheme, reaction):
    """WENO-Z7 component-wise (71) and characteristic (72) on 4 ghost layers."""
    print(f"[gpu] fr dq scheme {scheme} reaction {reaction}", flush=True)
    case = _case(tables, scheme=scheme, reaction=reaction, positive_preserving=1)
    for label, state in (("layer", None), ("premixed", _mixed_state(case, tables, 41 + scheme))):
        o, s = _pair(case, tables) if state is None else _pair_state(case, tables, *state)
        assert s.ngg == 4 and o.ngg == 4
        dq_o = o.rhs(0, 2e-9).copy()
        s.compute_rhs(0, 2e-9)
        err = rel_err(s.get(0, MX.FIELD_DQ), dq_o)
        print(f"  {label} dq per-variable:", " ".join(f"{e:.1e}" for e in err), flush=True)
        assert np.all(err[:5] <= 1e-12)
        assert np.all(err[5:] <= (1e-10 if reaction else 1e-12))
        s.close()


@pytest.mark.parametrize("scheme", [71, 72], ids=["wenoz7-cp", "wenoz7-char"])
def test_steps_weno7(tables, scheme):
    print(f"[gpu] fr 10 steps scheme {scheme}, then 5 steps on a 2 x 2 x 2 partition", flush=True)
    o, s = _pair(_case(tables, scheme=scheme, positive_preserving=1), tables)
    o.step(10)
    s.rk3(10, MX.MODE_GRAPH)
    _compare_state(o, s, 1, 1e-9, f"scheme {scheme}, 10 steps")
    s.close()
    case = pkg.cases.reacting_case([7, 6], [8, 9], [5, 6], tables[0], scheme=scheme)
    o, s = _pair(case, tables)
    o.step(5)
    s.rk3(5, MX.MODE_GRAPH)
    _compare_state(o, s, len(case["blocks"]), 1e-9, f"scheme {scheme}, 2x2x2, 5 steps")
    s.close()


def test_residual_matches_oracle(tables):
    """ostc_fr_store_last_step / ostc_fr_residual on identical states (the
    initial layer, then the oracle's state two steps later): only the
    summation order differs."""
    print("[gpu] fr residual", flush=True)
    case = pkg.cases.reacting_case([7, 6], [8, 9], [5, 6], tables[0], scheme=52)
    o, s = _pair(case, tables)
    nb = len(case["blocks"])
    for b in range(nb):
        s.set(b, MX.FIELD_BV, o.array(b, OF.BV))
    s.store_last_step()
    o.store_last_step()
    o.step(2)
    for b in range(nb):
        s.set(b, MX.FIELD_BV, o.array(b, OF.BV))
    got, want = s.compute_residual(), o.residual()
    print(f"  residual gpu {got} oracle {want}", flush=True)
    np.testing.assert_allclose(got, want, rtol=1e-12)
    assert np.all(want > 0)
    s.close()


@pytest.mark.parametrize("pp", [0, 1])
def test_stage_residual_reference_stencil(tables, pp):
    """Scheme 52 with ch_stencil 1: the reference's stencil of the five
    acoustic/entropy fields (src/WENO.cu:880-881) on the GPU and in the oracle."""
    print(f"[gpu] fr dq scheme 52, reference stencil, pp {pp}", flush=True)
    case = _case(tables, scheme=52, positive_preserving=pp)
    case["ch_stencil"] = 1
    o, s = _pair(case, tables)
    dt = 2e-9
    dq_o = o.rhs(0, dt).copy()
    s.compute_rhs(0, dt)
    err = rel_err(s.get(0, MX.FIELD_DQ), dq_o)
    print("  per-variable:", " ".join(f"{e:.1e}" for e in err), flush=True)
    assert np.all(err[:5] <= 1e-12)
    assert np.all(err[5:] <= 1e-10)
    s.close()


def test_steps_reference_stencil(tables):
    print("[gpu] fr 10 graph steps, scheme 52 with the reference stencil", flush=True)
    case = _case(tables, scheme=52, positive_preserving=1)
    case["ch_stencil"] = 1
    o, s = _pair(case, tables)
    dt_o = o.step(10)
    dt_g = s.rk3(10, MX.MODE_GRAPH)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o
    _compare_state(o, s, 1, 1e-9, "reference stencil, 10 steps")
    s.close()


def test_roe_entropy_fix_factor(tables):
    """entropy_fix_factor reaches the mixture Roe kernel (MUSCL, factor 0.6)."""
    print("[gpu] fr Roe dq, entropy_fix_factor 0.6", flush=True)
    case = _case(tables, scheme=2)
    case["entropy_fix_factor"] = 0.6
    o, s = _pair(case, tables)
    dq_o = o.rhs(0, 0.0).copy()
    s.compute_rhs(0, 0.0)
    err = rel_err(s.get(0, MX.FIELD_DQ), dq_o)
    print("  per-variable:", " ".join(f"{e:.1e}" for e in err), flush=True)
    assert np.all(err[:5] <= 1e-12)
    assert np.all(err[5:] <= 1e-10)
    s.close()
