"""The allocation-free momentum thickness (needs a GPU).

ostc_momentum_thickness_noalloc keeps its row sums in the library's static
device scratch instead of memory the context allocates.  Through the C ABI,
two calls in a row -- a 20^3 layer, then a smaller 12^3 one in another
context on the same GPU, so the second call finds the first one's values in
the scratch -- must give the value of ostc_momentum_thickness bit for bit.
A block with more rows than the scratch holds is refused with an error.
"""
import ctypes as C

import numpy as np
import pytest

from _util import cases, pkg

pytestmark = pytest.mark.gpu


def _theta_pair(n, steps):
    case = cases.fourier_case(n, n_steps=steps)
    s = pkg.Solver(case)
    s.setup_fourier_case()
    s.rk3(steps, pkg.MODE_BATCHED, want_dt=False)
    th = C.c_double()
    t = pkg.theta_struct(case["theta"])
    # the allocation-free call first, before anything of the objective is allocated
    rc = pkg.lib().ostc_momentum_thickness_noalloc(s.ctx, 0, C.byref(t), C.byref(th))
    assert rc == 0, pkg.lib().ostc_last_error(s.ctx)
    ref = s.momentum_thickness(0, case["theta"])
    s.close()
    return th.value, ref


def test_two_calls_match_the_allocating_entry_point():
    for n, steps in ((20, 3), (12, 2)):
        print(f"[gpu] momentum thickness without allocation, {n}^3 after {steps} steps", flush=True)
        got, ref = _theta_pair(n, steps)
        print(f"  noalloc {got:.17e} allocating {ref:.17e}", flush=True)
        assert np.isfinite(ref) and ref > 0
        assert got == ref


def test_block_taller_than_the_scratch_is_refused():
    print("[gpu] momentum thickness without allocation: a block of 16386 rows is refused", flush=True)
    case = cases.mixing_layer_case([8], [16385], [8], perturbation=0.0)
    blk = case["blocks"][0]
    theta = dict(u1=1.0, u2=-1.0, rho0=1.0, du=2.0, dy=case["dy"], ni=blk["mx"] - 1, nk=blk["mz"] - 1)
    s = pkg.Solver(case)
    print(f"  block {blk['mx']} x {blk['my']} x {blk['mz']} nodes", flush=True)
    try:
        s.momentum_thickness_noalloc(0, theta)
        msg = None
    except pkg.OstcError as e:
        msg = str(e)
    print(f"  error: {msg}", flush=True)
    s.close()
    assert msg is not None and "16386 rows" in msg and "covers at most 16384" in msg, msg
