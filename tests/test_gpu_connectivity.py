"""Connectivity records on the GPU (needs a GPU): the rotated two-block case
(block 1's index axes permuted and reversed, faces given as boundary N.txt
ranges and inner N.txt line pairs, an interface in two pieces) against the
oracle on the same records.

Tolerances (per variable, max|gpu - oracle| / max|oracle|): primitives (and
mass fractions) after n RK3 steps <= 1e-9, as for the aligned partitions.
The ABI refuses malformed records with OSTC_ERR_ARG and boundary labels
without an implementation (here 2, the wall) with OSTC_ERR_UNSUPPORTED.
"""
import ctypes as C

import numpy as np
import pytest

from _util import O, cases, oracle_pair, pkg, rel_err

pytestmark = pytest.mark.gpu

BV_TOL = 1e-9


def _global_bv(single):
    o = O.Oracle(single)
    o.setup_mixing_layer()
    return o.array(0, O.BV).copy()


def _solver_pair(pair, bv_global):
    s = pkg.Solver(pair)
    g = s.ngg
    for b, blk in enumerate(pair["blocks"]):
        jac, met = pkg.metric_from_coordinates(blk["mx"], blk["my"], blk["mz"], g,
                                               *cases.block_coordinates(pair, b, g))
        s.set_geometry(b, jac, met)
        s.set(b, pkg.FIELD_BV, cases.block_from_global(pair, b, bv_global, g))
    s.initialize()
    return s


@pytest.mark.parametrize("scheme,y_bc,mode", [
    (51, cases.BC_OUTFLOW, "plain"), (51, cases.BC_OUTFLOW, "batched"), (51, cases.BC_FARFIELD, "batched"),
    (71, cases.BC_OUTFLOW, "batched"), (52, cases.BC_FARFIELD, "plain"), (3, cases.BC_OUTFLOW, "plain")])
def test_rotated_pair_steps(scheme, y_bc, mode):
    n = 6
    print(f"[gpu] rotated pair, scheme {scheme}, y_bc {y_bc}, {mode}, {n} steps", flush=True)
    single = cases.mixing_layer_case([23], [15], [13], y_bc=y_bc, scheme=scheme)
    pair = cases.rotated_pair(single, 10, split_k=7)
    bvg = _global_bv(single)
    o = oracle_pair(pair, bvg)
    s = _solver_pair(pair, bvg)
    g = s.ngg
    dt_o = o.step(n)
    dt_g = s.rk3(n, pkg.MODE_PLAIN if mode == "plain" else pkg.MODE_BATCHED)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (dt_g, dt_o)
    for b in range(2):
        e = rel_err(s.get(b, pkg.FIELD_BV), o.array(b, O.BV))
        print(f"  block {b} bv rel err {e}", flush=True)
        assert np.all(e <= BV_TOL), (b, e)
    s.close()


def test_mixture_rotated_pair_steps():
    import mech
    import oracle_fr as OF
    n = 4
    print(f"[gpu] mixture rotated pair, scheme 52, farfield y, {n} steps", flush=True)
    T = mech.load(*pkg.mixture.H2_AIR_FILES)
    single = cases.reacting_case([19], [13], [11], T, scheme=52, y_bc=cases.BC_FARFIELD)
    pair = cases.rotated_pair(single, 8, split_k=6)
    o1 = OF.FrOracle(single, T, n_threads=8)
    o1.setup_mixing_layer()
    g = o1.ngg
    bvg, svg = o1.array(0, OF.BV).copy(), o1.array(0, OF.SV).copy()
    o = OF.FrOracle(pair, T, n_threads=8)
    s = pkg.mixture.FrSolver(pair, pkg.mixture.load_mechanism(*pkg.mixture.H2_AIR_FILES))
    for b, blk in enumerate(pair["blocks"]):
        jac, met = pkg.metric_from_coordinates(blk["mx"], blk["my"], blk["mz"], g,
                                               *cases.block_coordinates(pair, b, g))
        bv = cases.block_from_global(pair, b, bvg, g)
        sv = cases.block_from_global(pair, b, svg, g)
        o.array(b, OF.JAC)[...] = jac
        o.array(b, OF.METRIC)[...] = met
        o.array(b, OF.BV)[...] = bv
        o.array(b, OF.SV)[...] = sv
        s.set_geometry(b, jac, met)
        s.set(b, pkg.mixture.FIELD_BV, bv)
        s.set(b, pkg.mixture.FIELD_SV, sv)
    o.initialize()
    s.initialize()
    dt_o = o.step(n)
    dt_g = s.rk3(n)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (dt_g, dt_o)
    for b in range(2):
        for f, w in ((pkg.mixture.FIELD_BV, OF.BV), (pkg.mixture.FIELD_SV, OF.SV)):
            e = rel_err(s.get(b, f), o.array(b, w))
            print(f"  block {b} field {f} rel err {e}", flush=True)
            assert np.all(e <= BV_TOL), (b, f, e)
    s.close()


def test_malformed_records_are_refused():
    print("[gpu] connectivity record checks", flush=True)
    L = pkg.lib()
    single = cases.mixing_layer_case([15], [11], [9])
    pair = cases.rotated_pair(single, 6)
    s = pkg.Solver(dict(pair, boundaries=[], inner_faces=[]))
    ints = pkg.ints
    ctx = s.ctx
    nx, ny, nz = pair["global_shape"]
    ok_src, ok_tgt = pair["inner_faces"][1]["range"], pair["inner_faces"][1]["target_range"]
    # the finalized context takes no more faces
    assert L.ostc_add_inner_face_range(ctx, 0, ints(ok_src), ints(ok_tgt), 1) == -3
    s.close()
    # a context before finalize, for the record checks
    setup = pkg._Setup(5, 51, 2, 0, 0.0, 0.5, 0.72, 1.0, 1.0, 2)
    c = C.c_void_p()
    assert L.ostc_create(C.byref(setup), None, C.byref(c)) == 0
    for blk in pair["blocks"]:
        assert L.ostc_add_block(c, C.byref(pkg._BlockDesc(blk["mx"], blk["my"], blk["mz"])), None) == 0
    cut = pair["blocks"][0]["mx"] - 1
    nq = nx - cut
    bad = [
        ([cut + 1, cut + 1, 1, ny, 1, nz], ok_tgt),                  # no marked tangential axis
        ([cut + 1, cut + 1, -1, -ny, 1, nz], [-ny, -2, nq, nq, nz, 1]),  # lengths differ
        ([cut, cut, -1, -ny, 1, nz], ok_tgt),                       # normal index not on a face
        ([cut + 1, cut + 1, -1, -(ny + 1), 1, nz], [-(ny + 1), -1, 1, 1, nz, 1]),  # outside the block
    ]
    for src, tgt in bad:
        assert L.ostc_add_inner_face_range(c, 0, ints(src), ints(tgt), 1) == -1, (src, tgt)
    assert L.ostc_add_inner_face_range(c, 0, ints(ok_src), ints(ok_tgt), 1) == 0
    assert L.ostc_add_boundary_range(c, 0, ints([0, cut, 0, 0, 0, nz - 1]), 2, None) == -5  # wall
    assert "label 2" in L.ostc_last_error(c).decode()
    assert L.ostc_add_boundary_range(c, 0, ints([0, cut, 0, 0, 0, nz - 1]), 5, None) == -1  # inflow needs its state
    assert "ostc_add_inflow" in L.ostc_last_error(c).decode()
    assert L.ostc_add_boundary_range(c, 0, ints([0, cut, 3, 3, 0, nz - 1]), 6, None) == -1
    assert L.ostc_add_boundary_range(c, 0, ints([0, cut + 1, 0, 0, 0, nz - 1]), 6, None) == -1
    assert L.ostc_add_boundary_range(c, 0, ints([0, cut, 0, 0, 0, nz - 1]), 4, None) == -1  # farfield needs a state
    assert L.ostc_add_boundary_range(c, 0, ints([cut, 0, 0, 0, nz - 1, 0]), 6, None) == 0  # either order
    assert L.ostc_destroy(c) == 0
