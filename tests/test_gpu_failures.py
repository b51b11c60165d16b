"""Failure injection through the C ABI (needs a GPU).

The reference stops when the residual turns NaN and prints the nodes whose
primitives hold one (compute_residual and check_nan, src/Residual.cuh:165-184,
src/Residual.cu:35-49).  Here the residual entry points return
OSTC_ERR_NUMERIC with the block and node of the first NaN, and nothing aborts:
* a NaN poked into one node's primitives is found and located;
* a non-positive density poked into one node turns the state NaN within a
  step (the speed of sound of a negative temperature), on the per-block and
  on the batched path, and the residual reports it;
* the same for a mixture context.
No kernel faults on these states: NaN arithmetic runs to completion.
"""
import numpy as np
import pytest

from _util import cases, pkg

pytestmark = pytest.mark.gpu


def _solver(case):
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    s.initialize()
    return s


def test_nan_in_one_node_is_located():
    print("[gpu] NaN poked into block 1 node (3, 4, 2)", flush=True)
    case = cases.small_multiblock_case()
    s = _solver(case)
    s.store_last_step()
    s.rk3(1, pkg.MODE_BATCHED)
    g = s.ngg
    bv = s.get(1, pkg.FIELD_BV)
    bv[4, g + 2, g + 4, g + 3] = np.nan
    s.set(1, pkg.FIELD_BV, bv)
    with pytest.raises(pkg.OstcError, match=r"NaN in the residual: block 1, node \(3, 4, 2\)"):
        s.compute_residual()
    s.close()


@pytest.mark.parametrize("mode", [pkg.MODE_PLAIN, pkg.MODE_BATCHED], ids=["plain", "batched"])
def test_negative_density_is_reported(mode):
    print(f"[gpu] negative density poked, one step, mode {mode}", flush=True)
    case = cases.single_block_case(12)
    s = _solver(case)
    s.store_last_step()
    g = s.ngg
    bv = s.get(0, pkg.FIELD_BV)
    bv[0, g + 5, g + 5, g + 5] = -bv[0, g + 5, g + 5, g + 5]
    s.set(0, pkg.FIELD_BV, bv)
    s.initialize()  # cv and properties from the poked primitives, as a restart would
    s.rk3(1, mode, want_dt=False)
    with pytest.raises(pkg.OstcError, match="NaN in the residual: block 0"):
        s.compute_residual()
    s.close()


def test_mixture_nan_is_located():
    print("[gpu] NaN poked into a mixture node", flush=True)
    MX = pkg.mixture
    tg = MX.load_mechanism(*MX.H2_AIR_FILES)
    case = cases.reacting_single_block(10, 12, 8, tg, scheme=52)
    s = MX.FrSolver(case, tg)
    s.setup_mixing_layer()
    s.initialize()
    s.store_last_step()
    g = s.ngg
    bv = s.get(0, MX.FIELD_BV)
    bv[5, g + 1, g + 2, g + 6] = np.nan
    s.set(0, MX.FIELD_BV, bv)
    with pytest.raises(pkg.OstcError, match=r"NaN in the residual: block 0, node \(6, 2, 1\)"):
        s.compute_residual()
    s.close()
