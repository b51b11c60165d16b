"""The oracle of the exchange across ranks (oracle/parallel.py) pinned on the
CPU: record decoding worked by hand from ParallelFace::register_boundary
(src/Mesh.cpp:177-231), and what the exchange must achieve on a block cut in
two: after both sides pack and unpack, the two copies of the shared node
layer are equal (the mean of the two), and each side's ngg ghost layers hold
the other side's nodes beyond the shared layer, for either orientation of the
traversal; and the buffer layout itself, point by point from the bias
formula of setup_data_to_be_sent on a block whose primitives are their node
indices.  A planted error (layers packed outward, or a swapped loop order)
fails these.  The GPU kernels are compared with this oracle in
tests/test_gpu_parallel_faces.py.
"""
import numpy as np
import pytest

from _util import ROOT  # noqa: F401  (puts oracle/ on the path)
import parallel as P


def test_decode_low_face_by_hand():
    f = P.decode([1, 1, 1, 10, -1, -8], (5, 10, 8))
    assert f["face"] == 0 and f["direction"] == -1
    assert f["start"] == [0, 0, 0] and f["loop_dir"] == [0, 1, 1]
    assert f["n_point"] == [1, 10, 8] and f["loop_order"] == [0, 1, 2]


def test_decode_high_face_reversed_by_hand():
    f = P.decode([5, 5, -10, -1, 1, 8], (5, 10, 8))
    assert f["face"] == 0 and f["direction"] == 1
    assert f["start"] == [4, 9, 0] and f["loop_dir"] == [0, -1, 1]
    assert f["n_point"] == [1, 10, 8] and f["loop_order"] == [0, 2, 1]


def _split(seed, ngg=3, n=(11, 7, 6), cut=5):
    """A random padded block cut at node i = cut into A (0..cut) and B
    (cut..n-1), both padded with ngg ghost layers filled at random."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    full = rng.uniform(1, 2, size=(6, nz + 2 * ngg, ny + 2 * ngg, nx + 2 * ngg))
    A = rng.uniform(1, 2, size=(6, nz + 2 * ngg, ny + 2 * ngg, cut + 1 + 2 * ngg))
    B = rng.uniform(1, 2, size=(6, nz + 2 * ngg, ny + 2 * ngg, nx - cut + 2 * ngg))
    A[:, :, :, : cut + 1 + ngg] = full[:, :, :, : cut + 1 + ngg]
    B[:, :, :, ngg:] = full[:, :, :, cut + ngg:]
    # the two copies of the shared node layer differ before the exchange
    A[:, :, :, cut + ngg] += 0.25
    return full, A, B


@pytest.mark.parametrize("flip", [False, True], ids=["same-order", "reversed-y"])
def test_exchange_of_a_cut_block(flip):
    ngg, (nx, ny, nz), cut = 3, (11, 7, 6), 5
    full, A, B = _split(3, ngg, (nx, ny, nz), cut)
    a_shared = A[:, :, :, cut + ngg].copy()
    b_shared = B[:, :, :, ngg].copy()
    ma, mb = (cut + 1, ny, nz), (nx - cut, ny, nz)
    ya = [1, ny] if not flip else [-ny, -1]
    rec_a = [cut + 1, cut + 1, *ya, -1, -nz] if not flip else [cut + 1, cut + 1, ya[0], ya[1], 1, nz]
    rec_b = [1, 1, *ya, -1, -nz] if not flip else [1, 1, ya[0], ya[1], 1, nz]
    fa, fb = P.decode(rec_a, ma), P.decode(rec_b, mb)
    sa, sb = P.pack(A, fa, ngg), P.pack(B, fb, ngg)
    assert sa.size == sb.size == 6 * (ngg + 1) * ny * nz
    P.unpack(A, fa, ngg, sb)
    P.unpack(B, fb, ngg, sa)
    inner = (slice(None), slice(ngg, ngg + nz), slice(ngg, ngg + ny))
    mean = 0.5 * (a_shared + b_shared)[inner]
    np.testing.assert_array_equal(A[inner + (cut + ngg,)], mean)
    np.testing.assert_array_equal(B[inner + (ngg,)], mean)
    # A's ghosts beyond the cut hold B's nodes cut+1 .. cut+ngg (= the full block's)
    np.testing.assert_array_equal(A[inner + (slice(cut + 1 + ngg, cut + 1 + 2 * ngg),)],
                                  full[inner + (slice(cut + 1 + ngg, cut + 1 + 2 * ngg),)])
    # B's ghosts before the cut hold A's nodes cut-ngg .. cut-1
    np.testing.assert_array_equal(B[inner + (slice(0, ngg),)], full[inner + (slice(cut, cut + ngg),)])


def test_buffer_layout_by_hand():
    """Point t = n[loop_order[1]] * n_point[loop_order[2]] + n[loop_order[2]]
    holds the face node and then the nodes inward, nv values each."""
    ngg, m = 2, (5, 4, 3)
    bv = np.zeros((6,) + tuple(x + 2 * ngg for x in reversed(m)))
    K, J, I = np.meshgrid(*(np.arange(-ngg, x + ngg) for x in reversed(m)), indexing="ij")
    bv[0], bv[1], bv[2] = I, J, K
    # low x face; y traversed first (order 1), z negated: traversed fastest (order 2), from k = 2 down to 0
    f = P.decode([1, 1, 1, 4, -3, -1], m)
    buf = P.pack(bv, f, ngg).reshape(-1, ngg + 1, 6)
    assert buf.shape[0] == 4 * 3
    for t in range(12):
        j, k = t // 3, 2 - t % 3
        for layer in range(ngg + 1):
            assert tuple(buf[t, layer, :3]) == (layer, j, k), (t, layer, buf[t, layer, :3])
    # high y face, x negated and reversed (fastest), z first
    f = P.decode([-5, -1, 4, 4, 1, 3], m)
    buf = P.pack(bv, f, ngg).reshape(-1, ngg + 1, 6)
    for t in range(15):
        k, i = t // 5, 4 - t % 5
        for layer in range(ngg + 1):
            assert tuple(buf[t, layer, :3]) == (i, 3 - layer, k), (t, layer, buf[t, layer, :3])


def test_parallel_file_round_trip(tmp_path):
    """parallel N.txt as Mesh::read_parallel_interface reads it (src/Mesh.cpp:846-876):
    the counts line, a header line, then 9 integers per face in block order;
    mesh_io writes and reads it back into the records contexts register."""
    from _util import pkg
    mesh_io = pkg.mesh_io
    recs = [(0, [12, 12, 1, 9, -1, -7], 1, 11, 12), (1, [1, 1, -9, -1, 1, 7], 0, 21, 22),
            (1, [1, 5, 9, 9, -1, -7], 3, 31, 32)]
    path = tmp_path / mesh_io.file_name("parallel", 2)
    mesh_io.write_parallel(str(path), 2, recs)
    lines = path.read_text().splitlines()
    assert lines[0].split() == ["2", "3"] and lines[1].split() == ["1", "2"]
    assert mesh_io.read_parallel(str(path), 2) == recs
    got = mesh_io.parallel_faces(str(tmp_path), 2, rank=2)
    assert [(f["block"], f["range"], f["rank"], f["flag_send"], f["flag_recv"]) for f in got] == recs
