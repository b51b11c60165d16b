"""HIP kernels of the air schemes beyond component-wise WENO vs the CPU
oracle (needs a GPU): characteristic WENO-Z5 / WENO-Z7 (inviscid_scheme
52 / 72; per-block and fused batched kernels), the primitive reconstruction + Riemann path (1 LF, 3 AUSM+, 4 HLLC
with first-order, MUSCL or NND2 reconstruction) and the sixth-order
energy-preserving split form (6; reconstruction unused).

Tolerances (per variable, max|gpu - oracle| / max|oracle|):
  dq (inviscid + viscous)            <= 1e-12
  primitives after 10 RK3 steps      <= 1e-9
Both sides round every product and sum in the written order
(-ffp-contract=off); only sqrt and division are shared library routines.
The "rough" case perturbs the mixing layer by 5 % of du at Mc = 1.5, so the
minmod limiter takes both branches and AUSM+/HLLC see supersonic faces.
"""
import numpy as np
import pytest

from _util import O, cases, interior, oracle_for, pkg, rel_err

pytestmark = pytest.mark.gpu

DQ_TOL = 1e-12
BV_TOL = 1e-9

SCHEMES = [(52, 2), (72, 2), (1, 1), (1, 2), (2, 1), (2, 2), (3, 2), (3, 3), (4, 1), (4, 2), (4, 3), (6, 2)]
IDS = [f"s{s}-r{r}" for s, r in SCHEMES]


def _case(kind, scheme, recon):
    if kind == "single":
        return cases.single_block_case(20, scheme=scheme, reconstruction=recon)
    if kind == "multi":
        return cases.small_multiblock_case(scheme=scheme, reconstruction=recon)
    return cases.single_block_case(18, scheme=scheme, reconstruction=recon, ma_c=1.5, perturbation=5e-2)


@pytest.mark.parametrize("kind", ["single", "multi", "rough"])
@pytest.mark.parametrize("scheme,recon", SCHEMES, ids=IDS)
def test_rhs_dq(scheme, recon, kind):
    case = _case(kind, scheme, recon)
    print(f"[gpu] air scheme {scheme} recon {recon} rhs dq, {kind}", flush=True)
    o = oracle_for(case)
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    for b in range(len(case["blocks"])):
        s.set(b, pkg.FIELD_CV, o.array(b, O.CV))
        s.set(b, pkg.FIELD_BV, o.array(b, O.BV))
        s.set(b, pkg.FIELD_MUL, o.array(b, O.MUL))
    for b in range(len(case["blocks"])):
        o.rhs(b)
        s.compute_rhs(b)
        dq_g, dq_o = s.get(b, pkg.FIELD_DQ), o.array(b, O.DQ)
        e = rel_err(dq_g, dq_o)
        print(f"  block {b} dq rel err {e}", flush=True)
        assert np.all(e <= DQ_TOL), (b, e)
        if scheme in (2, 6):  # the Roe flux and EP6 are bit-identical to the oracle
            assert np.array_equal(dq_g, dq_o), (b, e)
    s.close()


@pytest.mark.parametrize("kind", ["single", "multi"])
@pytest.mark.parametrize("scheme,recon", SCHEMES, ids=IDS)
def test_steps(scheme, recon, kind):
    n = 10
    case = _case(kind, scheme, recon)
    print(f"[gpu] air scheme {scheme} recon {recon}, {n} plain RK3 steps, {kind}", flush=True)
    o = oracle_for(case)
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    s.initialize()
    dt_o = o.step(n)
    dt_g = s.rk3(n, pkg.MODE_PLAIN)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (dt_g, dt_o)
    g = s.ngg
    for b in range(len(case["blocks"])):
        e = rel_err(interior(s.get(b, pkg.FIELD_BV), g), interior(o.array(b, O.BV), g))
        print(f"  block {b} bv rel err {e}", flush=True)
        assert np.all(e <= BV_TOL), (b, e)
    s.close()


@pytest.mark.parametrize("factor", [0.0, 0.6])
def test_roe_entropy_fix_factor(factor):
    """entropy_fix_factor of the setup (src/Parameter.cpp:474) reaches the Roe
    kernels: dq at a non-default factor matches the oracle at that factor and
    differs from dq at the default 0.125 on the rough case."""
    print(f"[gpu] air Roe dq, entropy_fix_factor {factor}", flush=True)
    out = {}
    for f in (factor, 0.125):
        case = _case("rough", 2, 2)
        case["entropy_fix_factor"] = f
        o = oracle_for(case)
        s = pkg.Solver(case)
        s.setup_mixing_layer()
        s.set(0, pkg.FIELD_CV, o.array(0, O.CV))
        s.set(0, pkg.FIELD_BV, o.array(0, O.BV))
        s.set(0, pkg.FIELD_MUL, o.array(0, O.MUL))
        o.rhs(0)
        s.compute_rhs(0)
        out[f] = s.get(0, pkg.FIELD_DQ)
        e = rel_err(out[f], o.array(0, O.DQ))
        print(f"  factor {f}: dq rel err {e}", flush=True)
        assert np.all(e <= DQ_TOL), e
        s.close()
    assert np.max(np.abs(out[factor] - out[0.125])) > 0


@pytest.mark.parametrize("scheme", [52, 72])
@pytest.mark.parametrize("kind", ["single", "multi", "rough"])
@pytest.mark.parametrize("index", range(6))
def test_batched_characteristic_dq(scheme, kind, index):
    """The fused batched kernel's characteristic WENO-Z5 / Z7 (52 / 72) against
    the oracle, every tile shape: dq within the per-block path's 1e-12."""
    case = _case(kind, scheme, 2)
    print(f"[gpu] batched air scheme {scheme} dq, {kind}, tile entry {index}", flush=True)
    o = oracle_for(case)
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    for b in range(len(case["blocks"])):
        s.set(b, pkg.FIELD_CV, o.array(b, O.CV))
        s.set(b, pkg.FIELD_BV, o.array(b, O.BV))
        s.set(b, pkg.FIELD_MUL, o.array(b, O.MUL))
    s.set_tile(index, 0)
    for b in range(len(case["blocks"])):
        o.rhs(b)
        s.compute_rhs_batched(b)
        e = rel_err(s.get(b, pkg.FIELD_DQ), o.array(b, O.DQ))
        assert np.all(e <= DQ_TOL), (index, b, e)
    s.close()


@pytest.mark.parametrize("scheme", [52, 72])
@pytest.mark.parametrize("kind", ["single", "multi"])
def test_batched_characteristic_steps(scheme, kind):
    """10 batched RK3 steps (graph-replayed) with 52 / 72 against the oracle's
    sequential steps, and against the per-block path."""
    n = 10
    case = _case(kind, scheme, 2)
    print(f"[gpu] batched air scheme {scheme}, {n} RK3 steps, {kind}", flush=True)
    o = oracle_for(case)
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    s.initialize()
    dt_o = o.step(n)
    dt_g = s.rk3(n, pkg.MODE_BATCHED)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (dt_g, dt_o)
    g = s.ngg
    for b in range(len(case["blocks"])):
        e = rel_err(interior(s.get(b, pkg.FIELD_BV), g), interior(o.array(b, O.BV), g))
        print(f"  block {b} bv rel err {e}", flush=True)
        assert np.all(e <= BV_TOL), (b, e)
    s.close()


@pytest.mark.parametrize("scheme", [1, 2, 3, 4, 6])
def test_fused_paths_refuse_these_schemes(scheme):
    """The batched kernel implements schemes 50/51/71 and 52/72; the others
    are refused with an error instead of running another scheme."""
    case = cases.single_block_case(16, scheme=scheme, reconstruction=2)
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    s.initialize()
    with pytest.raises(pkg.OstcError, match="OSTC_MODE_PLAIN"):
        s.rk3(1, pkg.MODE_BATCHED)
    with pytest.raises(pkg.OstcError, match="OSTC_MODE_PLAIN"):
        s.compute_rhs_batched(0)
    s.close()
