"""Multi-block connectivity in the reference's file formats, on the CPU.

* the readers take the reference's own boundary / inner / parallel files
  (tests/golden/connectivity/, copies of input/boundary_condition/) and a
  write -> read round trip of the rotated two-block case;
* the oracle's inner-face relation (InnerFace::register_boundary,
  src/Mesh.cpp:38-176, and the loop of inner_communication,
  src/DataCommunication.cuh:60-105) delivers to every ghost node the global
  node index of the point it stands for -- a block whose index axes are
  permuted and reversed, faces split into pieces, and the periodic closure;
* the inviscid RHS is local along grid lines, so on the rotated two-block
  case the oracle reproduces the single-block run of the same box to
  round-off (a wrong orientation, offset or sign would be O(1) off);
* ostc_util_metric (host code of the product) equals the oracle's
  Block::compute_jac_metric on the rotated block's coordinates.
"""
import os
import tempfile

import numpy as np
import pytest

from _util import O, ROOT, cases, oracle_pair, pkg

mesh_io = pkg.mesh_io
FIX = os.path.join(ROOT, "tests", "golden", "connectivity")


def test_reads_the_reference_connectivity_files():
    bnd = mesh_io.read_boundary(os.path.join(FIX, mesh_io.file_name("boundary")), 1)
    assert [r[2] for r in bnd] == [8, 8, 5, 6, 6, 6]
    assert bnd[0] == (0, [0, 550, 0, 100, 0, 0], 8)
    assert bnd[3] == (0, [550, 550, 0, 100, 0, 80], 6)
    assert mesh_io.read_inner(os.path.join(FIX, mesh_io.file_name("inner")), 1) == []
    assert mesh_io.read_parallel(os.path.join(FIX, mesh_io.file_name("parallel")), 1) == []
    # label 5 (the mixing-layer inflow) needs its state
    with pytest.raises(ValueError, match="inflow label 5"):
        mesh_io.connectivity(FIX, 1)
    state = dict(inflow_type=2, delta_omega=1e-3, upper=[1, 2, 0, 0, 1e5, 300], lower=[1, -2, 0, 0, 1e5, 300])
    b, inner = mesh_io.connectivity(FIX, 1, inflows={5: state})
    assert len(b) == 6 and inner == []
    assert b[2] == dict(block=0, range=[0, 0, 0, 100, 0, 80], type=5, label=5, inflow=state)
    with pytest.raises(ValueError, match="label 8 has no implementation"):
        mesh_io.connectivity(FIX, 1, labels={5: 5, 6: 6}, inflows={5: state})


def _pair(y_bc=cases.BC_OUTFLOW, viscous_order=0, scheme=51):
    single = cases.mixing_layer_case([19], [13], [11], y_bc=y_bc, viscous_order=viscous_order, scheme=scheme)
    return single, cases.rotated_pair(single, 8, split_k=6)


def test_write_read_round_trip():
    _, pair = _pair()
    labels = [6 if f["type"] == cases.BC_OUTFLOW else 8 for f in pair["boundaries"]]
    with tempfile.TemporaryDirectory() as d:
        mesh_io.write_boundary(os.path.join(d, mesh_io.file_name("boundary")), 2,
                               [(f["block"], f["range"], lab) for f, lab in zip(pair["boundaries"], labels)])
        mesh_io.write_inner(os.path.join(d, mesh_io.file_name("inner")), 2,
                            [(f["block"], f["range"], f["target_range"], f["target"]) for f in pair["inner_faces"]])
        mesh_io.write_parallel(os.path.join(d, mesh_io.file_name("parallel")), 2)
        bnd, inner = mesh_io.connectivity(d, 2)
    assert [(f["block"], f["range"], f["type"]) for f in bnd] == \
        [(f["block"], f["range"], f["type"]) for f in pair["boundaries"]]
    assert inner == pair["inner_faces"]


def test_exchange_delivers_global_node_indices():
    _, pair = _pair()
    nx, ny, nz = pair["global_shape"]
    o = O.Oracle(pair)
    g = o.ngg
    sentinel = -1.0e30
    expect = []
    for b, blk in enumerate(pair["blocks"]):
        i, j, k = cases._local_grid(blk, g)
        want = np.stack([np.mod(i, nx - 1), j, k]).astype(np.float64)
        bv = o.array(b, O.BV)
        bv[...] = sentinel
        core = (slice(g, g + blk["mz"]), slice(g, g + blk["my"]), slice(g, g + blk["mx"]))
        for l in range(3):
            bv[l][core] = want[l][core]
        bv[3][core] = 0.0
        bv[4][core] = 1.0
        bv[5][core] = 1.0
        expect.append(want)
    assert O.lib().orc_communicate(o.ctx) == 0
    for b, blk in enumerate(pair["blocks"]):
        bv = o.array(b, O.BV)[:3]
        # the ghost layers of the inner faces: block 0 along i, block 1 along q
        a = 0 if b == 0 else 1
        m = blk["mx"] if b == 0 else blk["my"]
        sel = [slice(g, g + blk["mz"]), slice(g, g + blk["my"]), slice(g, g + blk["mx"])]
        for ghost in (slice(0, g), slice(g + m, 2 * g + m)):
            s = list(sel)
            s[2 - a] = ghost
            np.testing.assert_array_equal(bv[(slice(None), *s)], expect[b][(slice(None), *s)])
        # nothing else outside the interior was written
        mask = np.ones(bv.shape[1:], dtype=bool)
        mask[tuple(sel)] = False
        for ghost in (slice(0, g), slice(g + m, 2 * g + m)):
            s = list(sel)
            s[2 - a] = ghost
            mask[tuple(s)] = False
        assert np.all(bv[:, mask] == sentinel)


@pytest.mark.parametrize("scheme,y_bc", [(51, cases.BC_OUTFLOW), (52, cases.BC_FARFIELD), (71, cases.BC_OUTFLOW)])
def test_rotated_pair_reproduces_the_single_block(scheme, y_bc):
    single, pair = _pair(y_bc=y_bc, scheme=scheme)
    o1 = O.Oracle(single)
    o1.setup_mixing_layer()
    bv_global = o1.array(0, O.BV).copy()
    o1.initialize()
    o2 = oracle_pair(pair, bv_global)
    g = o1.ngg
    o1.step(4)
    o2.step(4)
    ref = o1.array(0, O.BV)
    for b in range(2):
        (k, j, i), got = cases.global_from_block(pair, b, o2.array(b, O.BV), g)
        want = ref[:, k + g, j + g, i + g]
        scale = np.max(np.abs(ref), axis=(1, 2, 3))
        err = np.max(np.abs(got - want), axis=1) / scale
        # round-off only: the rotated block sums its direction sweeps in
        # another order and runs the one-sided stencils mirrored
        assert np.all(err <= 1e-10), (b, err)


def test_product_metric_equals_the_oracle_metric():
    _, pair = _pair()
    o = O.Oracle(pair)
    g = o.ngg
    L = O.lib()
    for b, blk in enumerate(pair["blocks"]):
        xyz = cases.block_coordinates(pair, b, g)
        for which, arr in zip((O.X, O.Y, O.Z), xyz):
            o.array(b, which)[0][...] = arr
        assert L.orc_compute_metric(o.ctx, b) == 0
        jac, met = pkg.metric_from_coordinates(blk["mx"], blk["my"], blk["mz"], g, *xyz)
        np.testing.assert_array_equal(jac, o.array(b, O.JAC)[0])
        np.testing.assert_array_equal(met, o.array(b, O.METRIC))
    # the rotated block: xi along global -y, eta along -x, zeta along -z
    m = o.array(1, O.METRIC)[:, g, g, g]
    assert m[1] < 0 and m[3] < 0 and m[8] < 0 and m[0] == m[2] == m[4] == m[5] == m[6] == m[7] == 0


@pytest.mark.parametrize("scheme", [52, 71])
def test_mixture_rotated_pair_reproduces_the_single_block(scheme):
    """The finite-rate mixture oracle (reaction on, inviscid) on the rotated
    pair against its single-block run.  The first stage's dq agrees to
    round-off; the temperature Newton iteration (compute_temperature_and_pressure,
    src/FieldOperation.cu:3-40) stops at |1 - T_new/T| <= 1e-3, so a round-off
    difference can add or save one iteration at a node and move T by up to the
    square of that step: the bound is 1e-6, still far below the O(1) error of
    a wrong orientation."""
    import mech
    import oracle_fr as OF
    T = mech.load(*pkg.mixture.H2_AIR_FILES)
    single = cases.reacting_case([19], [13], [11], T, scheme=scheme, viscous_order=0)
    pair = cases.rotated_pair(single, 8, split_k=6)
    o1 = OF.FrOracle(single, T, n_threads=4)
    o1.setup_mixing_layer()
    g = o1.ngg
    bv_global, sv_global = o1.array(0, OF.BV).copy(), o1.array(0, OF.SV).copy()
    o1.initialize()
    o2 = OF.FrOracle(pair, T, n_threads=4)
    for b, blk in enumerate(pair["blocks"]):
        jac, met = pkg.metric_from_coordinates(blk["mx"], blk["my"], blk["mz"], g,
                                               *cases.block_coordinates(pair, b, g))
        o2.array(b, OF.JAC)[...] = jac
        o2.array(b, OF.METRIC)[...] = met
        o2.array(b, OF.BV)[...] = cases.block_from_global(pair, b, bv_global, g)
        o2.array(b, OF.SV)[...] = cases.block_from_global(pair, b, sv_global, g)
    o2.initialize()
    o1.step(2)
    o2.step(2)
    for which in (OF.BV, OF.SV):
        ref = o1.array(0, which)
        scale = np.max(np.abs(ref), axis=(1, 2, 3))
        for b in range(2):
            (k, j, i), got = cases.global_from_block(pair, b, o2.array(b, which), g)
            err = np.max(np.abs(got - ref[:, k + g, j + g, i + g]), axis=1) / scale
            assert np.all(err <= 1e-6), (which, b, err)
