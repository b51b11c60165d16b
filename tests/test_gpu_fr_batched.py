"""The batched step of the finite-rate path (needs a GPU).

OSTC_FR_MODE_BATCHED runs each kernel kind of a stage (source, the three
inviscid and three viscous sweeps, update, properties, local dt) as one
launch over a work-list of every block's tiles, captured in a graph; the
per-block modes launch each kernel once per block (the reference's loop over
blocks, src/RK.cuh:112-174).  Both run the same arithmetic in the same order
within a block, so:
* on uneven 2 x 2 x 2 partitions the batched state after a few steps equals
  the per-block state bit for bit (schemes 52, 51 + chemistry, 72, AUSM+);
* the batched state matches the oracle within 1e-9 after 5 steps;
* the timed batched mode reports one launch per kernel kind per stage.
"""
import numpy as np
import pytest

from _util import pkg, rel_err
import mech
import oracle_fr as OF

pytestmark = pytest.mark.gpu

MX = pkg.mixture
FILES = MX.H2_AIR_FILES
FIELDS = (MX.FIELD_CV, MX.FIELD_BV, MX.FIELD_SV, MX.FIELD_MUL)


@pytest.fixture(scope="module")
def tables():
    return MX.load_mechanism(*FILES), mech.load(*FILES)


def _solver(case, tg):
    s = MX.FrSolver(case, tg)
    s.setup_mixing_layer()
    s.initialize()
    return s


@pytest.mark.parametrize("scheme,kw", [(52, {}), (51, dict(positive_preserving=1)), (72, {}),
                                       (3, {})], ids=["52", "51-pp", "72", "ausm-muscl"])
def test_batched_equals_per_block(tables, scheme, kw):
    tg, _ = tables
    case = pkg.cases.reacting_case([7, 6], [8, 9], [5, 6], tg, scheme=scheme, **kw)
    print(f"[gpu] fr batched vs per-block, scheme {scheme}, {len(case['blocks'])} blocks, 3 steps", flush=True)
    a = _solver(case, tg)
    b = _solver(case, tg)
    dt_a = a.rk3(3, MX.MODE_GRAPH)
    dt_b = b.rk3(3, MX.MODE_BATCHED)
    assert dt_a == dt_b
    for blk in range(len(case["blocks"])):
        for f in FIELDS:
            np.testing.assert_array_equal(a.get(blk, f), b.get(blk, f), err_msg=f"block {blk} field {f}")
    a.close()
    b.close()


def test_batched_matches_oracle(tables):
    tg, to = tables
    case = pkg.cases.reacting_case([9, 5], [7, 10], [6, 4], tg, scheme=52)
    print("[gpu] fr batched 5 steps on an uneven 2 x 2 x 2 partition vs the oracle", flush=True)
    o = OF.FrOracle(case, to)
    o.setup_mixing_layer()
    o.initialize()
    s = _solver(case, tg)
    dt_o = o.step(5)
    dt_g = s.rk3(5, MX.MODE_BATCHED)
    print(f"  dt oracle {dt_o:.15e} gpu {dt_g:.15e}", flush=True)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o
    g = s.ngg
    for blk in range(len(case["blocks"])):
        for fo, fg in ((OF.BV, MX.FIELD_BV), (OF.SV, MX.FIELD_SV)):
            ig = s.get(blk, fg)[:, g:-g, g:-g, g:-g]
            io = o.array(blk, fo)[:, g:-g, g:-g, g:-g]
            e = rel_err(ig, io)
            assert np.all(e <= 1e-9), (blk, fo, e)
    s.close()


def test_batched_timed_launch_counts(tables):
    tg, _ = tables
    case = pkg.cases.reacting_case([7, 6], [8, 9], [5, 6], tg, scheme=52)
    print("[gpu] fr batched timed mode: one launch per kernel kind and stage", flush=True)
    s = _solver(case, tg)
    s.rk3(1, MX.MODE_BATCHED_TIMED)
    kms = s.kernel_ms()
    for kind, (ms, count) in kms.items():
        assert count == 3, (kind, count)
        assert ms > 0, kind
    s.rk3(1, MX.MODE_TIMED)
    per_block = s.kernel_ms()
    assert per_block["inviscid_x"][1] == 3 * len(case["blocks"])
    s.close()
