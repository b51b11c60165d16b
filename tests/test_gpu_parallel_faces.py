"""Faces shared with another rank on the GPU (needs a GPU): two contexts, one
block each, stand for two ranks of a block cut in two; each packs its
parallel face, the buffers are swapped on the host (the transfer), and each
unpacks the partner's.  The send buffers, and afterwards the primitives and
cv of both padded blocks, are bit-identical to the oracle of the reference's
setup_data_to_be_sent / assign_data_received (oracle/parallel.py, pinned in
tests/test_parallel_faces_host.py), for faces normal to each axis with
either tangential axis varying fastest, ranges forward and reversed, and
faces whose point count leaves a partial last workgroup.  Whole steps refuse
a context with parallel faces.
"""
import numpy as np
import pytest

from _util import ROOT, cases, pkg  # noqa: F401
import parallel as P

pytestmark = pytest.mark.gpu

GAMMA = 1.4


DIMS = (14, 11, 10)  # nodes of the cut domain
CUT = {0: 6, 1: 5, 2: 4}  # block A ends at node CUT + 1 (1-based) of the cut axis
# (cut axis, negated tangential axis, reverse the other tangential range,
# reverse the negated one): x-normal faces walk their layers, faces with x
# as the negated (fastest) axis walk their points, the rest neither.
ORIENT = [(0, 2, False, False), (0, 1, False, True), (1, 0, False, True), (1, 2, True, False),
          (2, 0, False, False), (2, 1, True, True)]


def _cells(axis, side):
    n = list(DIMS)
    n[axis] = CUT[axis] + 1 if side == 0 else DIMS[axis] - CUT[axis]
    return [[c - 1] for c in n]


def _record(axis, neg, rev_other, rev_neg, at):
    rec = []
    for d in range(3):
        if d == axis:
            rec += [at, at]
            continue
        lo, hi = 1, DIMS[d]
        if (rev_neg if d == neg else rev_other):
            lo, hi = hi, lo
        rec += [-lo, -hi] if d == neg else [lo, hi]
    return rec


def _case(cells, axis, face):
    c = cases.mixing_layer_case(*cells)
    c["boundaries"] = [b for b in c["boundaries"] if b.get("axis") != axis]
    c["parallel_faces"] = [face]
    return c


def _cv_of(bv):
    rho, u, v, w, p = bv[0], bv[1], bv[2], bv[3], bv[4]
    te = 0.5 * (u * u + v * v + w * w)
    te = te * rho
    te = te + p / (GAMMA - 1)
    return np.stack([rho, rho * u, rho * v, rho * w, te])


@pytest.mark.parametrize("axis,neg,rev_other,rev_neg", ORIENT, ids=[f"n{a}-neg{n}-r{int(r1)}{int(r2)}"
                                                                 for a, n, r1, r2 in ORIENT])
def test_two_ranks_exchange(axis, neg, rev_other, rev_neg):
    print(f"[gpu] parallel faces, two contexts, normal {axis}, negated {neg}, reversed {rev_other} {rev_neg}",
          flush=True)
    rec_a = _record(axis, neg, rev_other, rev_neg, CUT[axis] + 1)
    rec_b = _record(axis, neg, rev_other, rev_neg, 1)
    sa = pkg.Solver(_case(_cells(axis, 0), axis, dict(block=0, range=rec_a, rank=1, flag_send=7, flag_recv=8)))
    sb = pkg.Solver(_case(_cells(axis, 1), axis, dict(block=0, range=rec_b, rank=0, flag_send=8, flag_recv=7)))
    g = sa.ngg
    points = int(np.prod(DIMS)) // DIMS[axis]
    info = sa.parallel_faces()
    assert len(info) == 1 and info[0]["rank"] == 1 and info[0]["flag_send"] == 7
    assert info[0]["points"] == points and info[0]["count"] == 6 * (g + 1) * points
    rng = np.random.default_rng(11 + 4 * axis + neg)
    bva = rng.uniform(0.5, 2.0, size=sa.get(0, pkg.FIELD_BV).shape)
    bvb = rng.uniform(0.5, 2.0, size=sb.get(0, pkg.FIELD_BV).shape)
    sa.set(0, pkg.FIELD_BV, bva)
    sb.set(0, pkg.FIELD_BV, bvb)
    cva0, cvb0 = sa.get(0, pkg.FIELD_CV), sb.get(0, pkg.FIELD_CV)
    send_a, send_b = sa.pack_parallel(), sb.pack_parallel()
    na = [c[0] + 1 for c in _cells(axis, 0)]
    nb = [c[0] + 1 for c in _cells(axis, 1)]
    fa, fb = P.decode(rec_a, na), P.decode(rec_b, nb)
    np.testing.assert_array_equal(send_a, P.pack(bva, fa, g))
    np.testing.assert_array_equal(send_b, P.pack(bvb, fb, g))
    sa.unpack_parallel(send_b)
    sb.unpack_parallel(send_a)
    for s, bv0, cv0, f, recv in ((sa, bva, cva0, fa, send_b), (sb, bvb, cvb0, fb, send_a)):
        want = P.unpack(bv0.copy(), f, g, recv)
        got = s.get(0, pkg.FIELD_BV)
        np.testing.assert_array_equal(got, want)
        changed = np.any(want != bv0, axis=0)
        assert changed.sum() == (g + 1) * points
        cv = s.get(0, pkg.FIELD_CV)
        np.testing.assert_array_equal(cv[:, changed], _cv_of(want)[:, changed])
        np.testing.assert_array_equal(cv[:, ~changed], cv0[:, ~changed])
    with pytest.raises(pkg.OstcError, match="parallel faces"):
        sa.rk3(1, pkg.MODE_BATCHED)
    with pytest.raises(ValueError):
        sa.unpack_parallel(np.zeros(3))
    with pytest.raises(pkg.OstcError, match="pack"):  # unpack needs a pack since the last unpack
        sa.unpack_parallel(send_b)
    sa.close()
    sb.close()


def test_large_face_exchange():
    """A 130 x 70 face (9100 points: 285 workgroups of 32 points, the last one
    partial), y varying fastest and reversed: both send buffers and both
    unpacked blocks bit-identical to the oracle."""
    ny, nz, ca, cb = 130, 70, 6, 7
    print(f"[gpu] parallel faces, large face {ny} x {nz}", flush=True)
    rec_a = [ca, ca, -ny, -1, 1, nz]
    rec_b = [1, 1, -ny, -1, 1, nz]
    cut = lambda x: cases.mixing_layer_case([x - 1], [ny - 1], [nz - 1])
    solvers, faces, states = [], [], []
    for nx, rec, rank in ((ca, rec_a, 1), (cb, rec_b, 0)):
        c = cut(nx)
        c["boundaries"] = [b for b in c["boundaries"] if b.get("axis") != 0]
        c["parallel_faces"] = [dict(block=0, range=rec, rank=rank)]
        s = pkg.Solver(c)
        bv = np.random.default_rng(40 + nx).uniform(0.5, 2.0, size=s.get(0, pkg.FIELD_BV).shape)
        s.set(0, pkg.FIELD_BV, bv)
        solvers.append(s)
        faces.append(P.decode(rec, (nx, ny, nz)))
        states.append(bv)
    g = solvers[0].ngg
    sends = [s.pack_parallel() for s in solvers]
    for s, f, bv, send in zip(solvers, faces, states, sends):
        np.testing.assert_array_equal(send, P.pack(bv, f, g))
    solvers[0].unpack_parallel(sends[1])
    solvers[1].unpack_parallel(sends[0])
    for s, f, bv, recv in zip(solvers, faces, states, sends[::-1]):
        np.testing.assert_array_equal(s.get(0, pkg.FIELD_BV), P.unpack(bv.copy(), f, g, recv))
        s.close()


# Two parallel faces of one block that share the edge line i = mx-1, j = my-1:
# the second face's unpack averages with the first face's means there, as the
# reference's in-order assign_data_received does.
EDGE_M = (8, 9, 7)
EDGE_RECS = ([8, 8, 1, 9, -1, -7], [-1, -8, 9, 9, 1, 7])


def _edge_case(make, m):
    c = make([[m[0] - 1], [m[1] - 1], [m[2] - 1]])
    c["boundaries"] = [b for b in c["boundaries"] if b.get("axis") == 2]
    c["parallel_faces"] = [dict(block=0, range=r, rank=1 + q) for q, r in enumerate(EDGE_RECS)]
    return c


def test_faces_sharing_an_edge():
    print("[gpu] parallel faces sharing an edge, air", flush=True)
    s = pkg.Solver(_edge_case(lambda cells: cases.mixing_layer_case(*cells), EDGE_M))
    g = s.ngg
    info = s.parallel_faces()
    assert [f["points"] for f in info] == [9 * 7, 8 * 7]
    rng = np.random.default_rng(21)
    bv = rng.uniform(0.5, 2.0, size=s.get(0, pkg.FIELD_BV).shape)
    s.set(0, pkg.FIELD_BV, bv)
    send = s.pack_parallel()
    faces = [P.decode(r, EDGE_M) for r in EDGE_RECS]
    recv = rng.uniform(0.5, 2.0, size=send.shape)
    want = bv.copy()
    for f, face in zip(info, faces):
        lo, hi = f["offset"], f["offset"] + f["count"]
        np.testing.assert_array_equal(send[lo:hi], P.pack(bv, face, g))
        P.unpack(want, face, g, recv[lo:hi])
    s.unpack_parallel(recv)
    np.testing.assert_array_equal(s.get(0, pkg.FIELD_BV), want)
    changed = np.any(want != bv, axis=0)
    np.testing.assert_array_equal(s.get(0, pkg.FIELD_CV)[:, changed], _cv_of(want)[:, changed])
    s.close()


def test_bad_records_are_refused():
    for rec, msg in (([1, 3, 1, 9, -1, -7], "no axis with equal ends"),
                     ([4, 4, 1, 9, -1, -7], "not on a block face"),
                     ([1, 1, 1, 9, 1, 7], "negated"),
                     ([1, 1, 1, 12, -1, -7], "outside the block")):
        with pytest.raises(pkg.OstcError, match=msg):
            pkg.Solver(_case([[5], [8], [6]], 0, dict(block=0, range=rec, rank=1)))


@pytest.fixture(scope="module")
def fr_tables():
    import mech
    MX = pkg.mixture
    return MX.load_mechanism(*MX.H2_AIR_FILES), mech.load(*MX.H2_AIR_FILES)


def _fr_case(tables, cells, axis, face):
    c = cases.reacting_case(*cells, tables)
    c["boundaries"] = [b for b in c["boundaries"] if b.get("axis") != axis]
    c["parallel_faces"] = [face]
    return c


def test_faces_sharing_an_edge_mixture(fr_tables):
    import oracle_fr as OF
    MX = pkg.mixture
    tg, to = fr_tables
    print("[gpu] parallel faces sharing an edge, mixture", flush=True)
    case = _edge_case(lambda cells: cases.reacting_case(*cells, tg), EDGE_M)
    s = MX.FrSolver(case, tg)
    g = s.ngg
    info = s.parallel_faces()
    rng = np.random.default_rng(22)
    bv, sv = _fr_state(rng, s.get(0, MX.FIELD_BV).shape, s.get(0, MX.FIELD_SV).shape)
    s.set(0, MX.FIELD_BV, bv)
    s.set(0, MX.FIELD_SV, sv)
    full = np.concatenate([bv, sv])
    send = s.pack_parallel()
    recv_bv, recv_sv = _fr_state(rng, (6, send.size), (s.ns, send.size))
    recv = send.copy().reshape(-1, 6 + s.ns)
    recv[:, :6] = recv_bv[:, :recv.shape[0]].T
    recv[:, 6:] = recv_sv[:, :recv.shape[0]].T
    recv = recv.ravel()
    want = full.copy()
    for f, r in zip(info, EDGE_RECS):
        lo, hi = f["offset"], f["offset"] + f["count"]
        P.unpack(want, P.decode(r, EDGE_M), g, recv[lo:hi])
    s.unpack_parallel(recv)
    np.testing.assert_array_equal(s.get(0, MX.FIELD_BV), want[:6])
    np.testing.assert_array_equal(s.get(0, MX.FIELD_SV), want[6:])
    o = OF.FrOracle(case, to)
    o.array(0, OF.BV)[...] = want[:6]
    o.array(0, OF.SV)[...] = want[6:]
    o.L.orf_cv_from_bv(o.ctx, 0)
    changed = np.any(want != full, axis=0)
    cv, cvo = s.get(0, MX.FIELD_CV), o.array(0, OF.CV)
    err = np.max(np.abs(cv[:, changed] - cvo[:, changed]), axis=1) / np.max(np.abs(cvo[:, changed]), axis=1)
    assert np.all(err <= 1e-13), err
    o.close()
    s.close()


def _fr_state(rng, shape_bv, shape_sv):
    bv = np.empty(shape_bv)
    bv[0] = rng.uniform(0.5, 2.0, shape_bv[1:])
    bv[1:4] = rng.uniform(-100, 100, (3,) + shape_bv[1:])
    bv[4] = rng.uniform(5e4, 2e5, shape_bv[1:])
    bv[5] = rng.uniform(300, 2000, shape_bv[1:])
    sv = rng.uniform(0.01, 1.0, shape_sv)
    sv /= sv.sum(axis=0, keepdims=True)
    return bv, sv


@pytest.mark.parametrize("axis,neg,rev_other,rev_neg", [ORIENT[0], ORIENT[2]], ids=["n0-neg2", "n1-neg0"])
def test_two_ranks_exchange_mixture(fr_tables, axis, neg, rev_other, rev_neg):
    """The same on two finite-rate contexts: 6 primitives and 9 mass
    fractions per layer; cv of the rewritten nodes against the mixture
    oracle's cv from primitives (NASA-7 energy) to 1e-13."""
    import oracle_fr as OF
    MX = pkg.mixture
    tg, to = fr_tables
    print(f"[gpu] parallel faces, two mixture contexts, normal {axis}, negated {neg}", flush=True)
    rec_a = _record(axis, neg, rev_other, rev_neg, CUT[axis] + 1)
    rec_b = _record(axis, neg, rev_other, rev_neg, 1)
    ca = _fr_case(tg, _cells(axis, 0), axis, dict(block=0, range=rec_a, rank=1))
    cb = _fr_case(tg, _cells(axis, 1), axis, dict(block=0, range=rec_b, rank=0))
    sa, sb = MX.FrSolver(ca, tg), MX.FrSolver(cb, tg)
    g = sa.ngg
    points = int(np.prod(DIMS)) // DIMS[axis]
    assert sa.parallel_faces()[0]["count"] == 15 * (g + 1) * points
    rng = np.random.default_rng(5)
    states = []
    for s in (sa, sb):
        bv, sv = _fr_state(rng, s.get(0, MX.FIELD_BV).shape, s.get(0, MX.FIELD_SV).shape)
        s.set(0, MX.FIELD_BV, bv)
        s.set(0, MX.FIELD_SV, sv)
        states.append((bv, sv, s.get(0, MX.FIELD_CV)))
    send_a, send_b = sa.pack_parallel(), sb.pack_parallel()
    fa = P.decode(rec_a, [c[0] + 1 for c in _cells(axis, 0)])
    fb = P.decode(rec_b, [c[0] + 1 for c in _cells(axis, 1)])
    both = [np.concatenate([bv, sv]) for bv, sv, _ in states]
    np.testing.assert_array_equal(send_a, P.pack(both[0], fa, g))
    np.testing.assert_array_equal(send_b, P.pack(both[1], fb, g))
    sa.unpack_parallel(send_b)
    sb.unpack_parallel(send_a)
    for s, case, full, (bv0, sv0, cv0), f, recv in ((sa, ca, both[0], states[0], fa, send_b),
                                                    (sb, cb, both[1], states[1], fb, send_a)):
        want = P.unpack(full.copy(), f, g, recv)
        np.testing.assert_array_equal(s.get(0, MX.FIELD_BV), want[:6])
        np.testing.assert_array_equal(s.get(0, MX.FIELD_SV), want[6:])
        o = OF.FrOracle(case, to)
        o.array(0, OF.BV)[...] = want[:6]
        o.array(0, OF.SV)[...] = want[6:]
        o.L.orf_cv_from_bv(o.ctx, 0)
        changed = np.any(want != full, axis=0)
        cv, cvo = s.get(0, MX.FIELD_CV), o.array(0, OF.CV)
        err = np.max(np.abs(cv[:, changed] - cvo[:, changed]), axis=1) / np.max(np.abs(cvo[:, changed]), axis=1)
        assert np.all(err <= 1e-13), err
        np.testing.assert_array_equal(cv[:, ~changed], cv0[:, ~changed])
        o.close()
    with pytest.raises(pkg.OstcError, match="parallel faces"):
        sa.rk3(1, MX.MODE_BATCHED)
    sa.close()
    sb.close()
