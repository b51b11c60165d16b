"""Dual time stepping with the DPLUR inner solver on the GPU against the
oracle (needs a GPU): the mixing layer after a few physical steps, single
block and a 3 x 2 x 2 partition, with the inner-iteration monitor on and
off; primitives within 1e-9 and the same inner-iteration counts.  The
oracle's restatement is pinned in tests/test_oracle_dual_time.py (flux
Jacobian against finite differences, BDF2 at convergence, free stream).
"""
import numpy as np
import pytest

from _util import O, cases, interior, oracle_for, pkg, rel_err

pytestmark = pytest.mark.gpu


def _run(case, settings, steps, rk_steps=1):
    o = oracle_for(case)
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    s.initialize()
    o.step(rk_steps)
    s.rk3(rk_steps, pkg.MODE_PLAIN)
    o.dual_time_init(settings)
    s.set_dual_time(settings)
    counts_o, counts_g = [], []
    for _ in range(steps):
        counts_o.append(o.dual_time_step(1))
        counts_g.append(s.dual_time_step(1))
    return o, s, counts_o, counts_g


@pytest.mark.parametrize("adjust", [0, 2], ids=["fixed", "monitored"])
def test_single_block(adjust):
    print(f"[gpu] dual time stepping, 12^3 block, adjust {adjust}", flush=True)
    case = cases.single_block_case(12)
    settings = dict(dt=4e-8, inner_iteration=6, max_inner_iteration=20, iteration_adjust_step=adjust,
                    dplur_inner_step=3, step=0)
    o, s, co, cg = _run(case, settings, 4)
    print(f"  inner iterations oracle {co} gpu {cg}", flush=True)
    assert co == cg
    e = rel_err(interior(s.get(0, pkg.FIELD_BV), s.ngg), interior(o.array(0, O.BV), o.ngg))
    print(f"  bv rel err {e}", flush=True)
    assert np.all(e <= 1e-9), e
    s.close()


def test_multiblock_partition():
    case = cases.small_multiblock_case()
    print(f"[gpu] dual time stepping, {len(case['blocks'])} blocks", flush=True)
    settings = dict(dt=3e-8, inner_iteration=5, max_inner_iteration=20, iteration_adjust_step=3,
                    dplur_inner_step=2, step=0)
    o, s, co, cg = _run(case, settings, 3)
    assert co == cg
    worst = 0.0
    for b in range(len(case["blocks"])):
        e = rel_err(interior(s.get(b, pkg.FIELD_BV), s.ngg), interior(o.array(b, O.BV), o.ngg))
        worst = max(worst, float(e.max()))
        assert np.all(e <= 1e-9), (b, e)
    print(f"  worst bv rel err {worst:.2e}", flush=True)
    s.close()


def test_error_paths():
    print("[gpu] dual time stepping error paths", flush=True)
    case = cases.single_block_case(10)
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    s.initialize()
    with pytest.raises(pkg.OstcError, match="ostc_set_dual_time"):
        s.dual_time_step(1)
    with pytest.raises(pkg.OstcError, match="bad dual time"):
        s.set_dual_time(dict(dt=-1.0))
    s.close()
