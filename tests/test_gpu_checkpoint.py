"""Restart from the reference's checkpoint format on the GPU (needs a GPU):
n steps, flowfield.plt with all ghost layers, a new context loaded from it,
n more steps, against 2n steps without a stop.  The restart rebuilds cv from
the primitives (as the reference's does), so the two runs differ by
round-off: <= 1e-10.  Without the ghost layers (the reference's default
output) the edge ghosts, which no exchange fills, restart from the initial
state instead."""
import os
import tempfile

import numpy as np
import pytest

from _util import cases, pkg, rel_err

pytestmark = pytest.mark.gpu


def test_air_restart():
    print("[gpu] air restart from flowfield.plt, 3 + 3 steps", flush=True)
    case = cases.small_multiblock_case()
    ref = pkg.Solver(case)
    ref.setup_mixing_layer()
    ref.initialize()
    ref.rk3(6, pkg.MODE_BATCHED)
    a = pkg.Solver(case)
    a.setup_mixing_layer()
    a.initialize()
    a.rk3(3, pkg.MODE_BATCHED)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "flowfield.plt")
        pkg.checkpoint.save(a, path, solution_time=1.0, ghosts=a.ngg)
        b = pkg.Solver(case)
        b.setup_mixing_layer()
        assert pkg.checkpoint.load(b, path) == 1.0
    b.rk3(3, pkg.MODE_BATCHED)
    for blk in range(len(case["blocks"])):
        e = rel_err(b.get(blk, pkg.FIELD_BV), ref.get(blk, pkg.FIELD_BV))
        print(f"  block {blk} bv rel err {e}", flush=True)
        assert np.all(e <= 1e-10), (blk, e)
    for s in (ref, a, b):
        s.close()


def test_mixture_restart():
    print("[gpu] mixture restart from flowfield.plt, 2 + 2 steps", flush=True)
    M = pkg.mixture
    tg = M.load_mechanism(*M.H2_AIR_FILES)
    case = cases.reacting_single_block(14, 16, 10, tg, scheme=52)

    def fresh():
        s = M.FrSolver(case, tg)
        s.setup_mixing_layer()
        s.initialize()
        return s
    ref = fresh()
    ref.rk3(4)
    a = fresh()
    a.rk3(2)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "flowfield.plt")
        pkg.checkpoint.save(a, path, ghosts=a.ngg)
        data = pkg.checkpoint.read_plt(path)
        assert data["names"][10:] == tg.species()
        b = M.FrSolver(case, tg)
        b.setup_mixing_layer()
        pkg.checkpoint.load(b, path)
    b.rk3(2)
    for f in (M.FIELD_BV, M.FIELD_SV):
        e = rel_err(b.get(0, f), ref.get(0, f))
        print(f"  field {f} rel err {e}", flush=True)
        assert np.all(e <= 1e-10), (f, e)
    for s in (ref, a, b):
        s.close()
