"""Restart from the reference's checkpoint format on the GPU (needs a GPU):
3 steps, flowfield.plt, a new context loaded from it and 3 more steps,
against the oracle restarted from the same file by the same procedure
(primitives of the nodes, then boundary conditions, exchange, properties and
cv): primitives (and mass fractions) <= 1e-9.

A restart does not continue a run bit for bit, in the reference either: the
boundary conditions run before the exchange in every stage, so the ghost
nodes on block edges hold values one stage old, and the restart's fresh
boundary pass replaces them."""
import os
import tempfile

import numpy as np
import pytest

from _util import O, cases, oracle_for, pkg, rel_err

pytestmark = pytest.mark.gpu


def _load_into(arr_bv, zone, g, names, arr_sv=None, species=()):
    keep = (zone["shape"][2] - (arr_bv.shape[-1] - 2 * g)) // 2
    c = g - keep
    sl = (slice(c, arr_bv.shape[-3] - c), slice(c, arr_bv.shape[-2] - c), slice(c, arr_bv.shape[-1] - c))
    for l, nm in enumerate(names):
        arr_bv[(l,) + sl] = zone[nm]
    for l, nm in enumerate(species):
        arr_sv[(l,) + sl] = zone[nm]


@pytest.mark.parametrize("ghosts", [0, 3])
def test_air_restart_matches_the_oracle_restart(ghosts):
    print(f"[gpu] air restart from flowfield.plt ({ghosts} ghost layers), 3 + 3 steps", flush=True)
    case = cases.small_multiblock_case()
    a = pkg.Solver(case)
    a.setup_mixing_layer()
    a.initialize()
    a.rk3(3, pkg.MODE_BATCHED)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "flowfield.plt")
        pkg.checkpoint.save(a, path, solution_time=1.0, ghosts=ghosts)
        b = pkg.Solver(case)
        b.setup_mixing_layer()
        assert pkg.checkpoint.load(b, path) == 1.0
        data = pkg.checkpoint.read_plt(path)
    o = oracle_for(case, steps_init=False)
    g = o.ngg
    for blk, zone in enumerate(data["zones"]):
        _load_into(o.array(blk, O.BV), zone, g, pkg.checkpoint.BASE_NAMES[3:9])
    o.initialize()
    b.rk3(3, pkg.MODE_BATCHED)
    o.step(3)
    for blk in range(len(case["blocks"])):
        e = rel_err(b.get(blk, pkg.FIELD_BV), o.array(blk, O.BV))
        print(f"  block {blk} bv rel err {e}", flush=True)
        assert np.all(e <= 1e-9), (blk, e)
    a.close()
    b.close()


def test_mixture_restart_matches_the_oracle_restart():
    import mech
    import oracle_fr as OF
    print("[gpu] mixture restart from flowfield.plt, 2 + 2 steps", flush=True)
    M = pkg.mixture
    tg = M.load_mechanism(*M.H2_AIR_FILES)
    to = mech.load(*M.H2_AIR_FILES)
    case = cases.reacting_single_block(14, 16, 10, tg, scheme=52)
    a = M.FrSolver(case, tg)
    a.setup_mixing_layer()
    a.initialize()
    a.rk3(2)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "flowfield.plt")
        pkg.checkpoint.save(a, path)
        data = pkg.checkpoint.read_plt(path)
        assert data["names"][10:] == tg.species()
        b = M.FrSolver(case, tg)
        b.setup_mixing_layer()
        pkg.checkpoint.load(b, path)
    o = OF.FrOracle(case, to, n_threads=8)
    o.setup_mixing_layer()
    _load_into(o.array(0, OF.BV), data["zones"][0], o.ngg, pkg.checkpoint.BASE_NAMES[3:9],
               o.array(0, OF.SV), tg.species())
    o.initialize()
    b.rk3(2)
    o.step(2)
    for f, w in ((M.FIELD_BV, OF.BV), (M.FIELD_SV, OF.SV)):
        e = rel_err(b.get(0, f), o.array(0, w))
        print(f"  field {f} rel err {e}", flush=True)
        assert np.all(e <= 1e-9), (f, e)
    a.close()
    b.close()
