"""The sponge layer of the mixture path on the GPU (needs a GPU), against
the oracle: the spatially developing reacting layer with sponges on x+, y-
and y+ after 4 RK3 steps (primitives and mass fractions <= 1e-9), and the
reference's unscaled term in one RHS evaluation (dq <= 1e-12; stepping with
it is unstable on a metre-scaled grid, whose Jacobian is ~1e-15); the air
path's sponge on the per-block kernels and in the fused batched kernel.
"""
import numpy as np
import pytest

from _util import cases, pkg, rel_err

pytestmark = pytest.mark.gpu


def _pair(jacobian_scaled):
    import mech
    import oracle_fr as OF
    tg = pkg.mixture.load_mechanism(*pkg.mixture.H2_AIR_FILES)
    to = mech.load(*pkg.mixture.H2_AIR_FILES)
    case = cases.reacting_case([15], [19], [9], to, spatial=True, inflow_noise=1e-2)
    case["sponge"] = cases.sponge_layer(case, sides=(1, 2, 3), depth=0.3, coefficient=2.0,
                                        jacobian_scaled=jacobian_scaled)
    o = OF.FrOracle(case, to, n_threads=8)
    o.setup_mixing_layer()
    o.initialize()
    s = pkg.mixture.FrSolver(case, tg)
    s.setup_mixing_layer()
    s.initialize()
    return OF, o, s


def test_sponge_steps():
    print("[gpu] mixture spatial layer with sponge, 4 steps", flush=True)
    OF, o, s = _pair(1)
    dt_o = o.step(4)
    dt_g = s.rk3(4)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (dt_g, dt_o)
    for f, w in ((pkg.mixture.FIELD_BV, OF.BV), (pkg.mixture.FIELD_SV, OF.SV)):
        e = rel_err(s.get(0, f), o.array(0, w))
        print(f"  field {f} rel err {e}", flush=True)
        assert np.all(e <= 1e-9), (f, e)
    s.close()


def test_unscaled_sponge_rhs():
    print("[gpu] mixture rhs with the unscaled sponge term", flush=True)
    OF, o, s = _pair(0)
    dt = 1e-9
    dq_o = o.rhs(0, dt).copy()
    s.compute_rhs(0, dt)
    dq_g = s.get(0, pkg.mixture.FIELD_DQ)
    e = rel_err(dq_g, dq_o)
    print(f"  dq rel err {e}", flush=True)
    assert np.all(e <= 1e-10), e   # the chemistry rows: <= 1e-10 (the source gate)
    s.close()


@pytest.mark.parametrize("mode", [pkg.MODE_PLAIN, pkg.MODE_BATCHED], ids=["plain", "batched"])
def test_air_sponge_steps(mode):
    """The air path with the sponge term: per-block kernels (PLAIN) and the
    fused batched kernel, which adds the term first in each stage's dq."""
    from _util import O, oracle_for
    print(f"[gpu] air spatial layer with sponge, mode {mode}, 6 steps", flush=True)
    case = cases.mixing_layer_case([19], [15], [11], spatial=True, inflow_noise=1e-2)
    case["sponge"] = cases.sponge_layer(case, sides=(1, 2, 3), depth=0.3, coefficient=2.0)
    o = oracle_for(case)
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    s.initialize()
    dt_o = o.step(6)
    dt_g = s.rk3(6, mode)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (dt_g, dt_o)
    e = rel_err(s.get(0, pkg.FIELD_BV), o.array(0, O.BV))
    print(f"  bv rel err {e}", flush=True)
    assert np.all(e <= 1e-9), e
    s.close()
