"""Oracle (test infrastructure only): the reference's exchange across ranks,
restated in numpy.  Product code never imports this module.

* ``decode`` -- ParallelFace::register_boundary (src/Mesh.cpp:177-231) for a
  three-dimensional block: the normal axis is the one with equal ends (the
  low face when that end is 1, direction -1), the tangential axis given
  negated is ``loop_order[2]`` (traversed fastest), the other ``loop_order[1]``;
  ranges become 0-based with ``loop_dir`` = +-1.
* ``pack`` -- setup_data_to_be_sent (src/DataCommunication.cu:3-40): per face
  point ``n[loop_order[1]] * n_point[loop_order[2]] + n[loop_order[2]]``, the
  face node and then ngg nodes inward, all primitives of each.
* ``unpack`` -- assign_data_received (src/DataCommunication.cuh:201-240): the
  face node takes 0.5 (own + received), the ngg nodes outward the received
  inner nodes (cv is recomputed from the primitives by the caller).

Arrays are the padded primitive planes bv[l, k + g, j + g, i + g].
"""
import numpy as np


def decode(rec, m):
    lo = [rec[0], rec[2], rec[4]]
    hi = [rec[1], rec[3], rec[5]]
    face = next(d for d in range(3) if lo[d] == hi[d])
    direction = -1 if lo[face] == 1 else 1
    start, end = [0, 0, 0], [0, 0, 0]
    start[face] = end[face] = lo[face] - 1
    neg = next(d for d in range(3) if lo[d] < 0)
    start[neg], end[neg] = -lo[neg] - 1, -hi[neg] - 1
    other = 3 - face - neg
    start[other], end[other] = lo[other] - 1, hi[other] - 1
    loop_dir = [0, 0, 0]
    for d in (neg, other):
        loop_dir[d] = 1 if start[d] < end[d] else -1
    n_point = [abs(start[d] - end[d]) + 1 for d in range(3)]
    return dict(face=face, direction=direction, start=start, loop_dir=loop_dir, n_point=n_point,
                loop_order=[face, other, neg])


def _points(f):
    lo1, lo2 = f["loop_order"][1], f["loop_order"][2]
    for n1 in range(f["n_point"][lo1]):
        for n2 in range(f["n_point"][lo2]):
            idx = list(f["start"])
            idx[lo1] += n1 * f["loop_dir"][lo1]
            idx[lo2] += n2 * f["loop_dir"][lo2]
            yield idx


def pack(bv, f, ngg):
    nv = bv.shape[0]
    out = []
    for idx in _points(f):
        for layer in range(ngg + 1):
            i = list(idx)
            i[f["face"]] -= layer * f["direction"]
            out.append(bv[:, i[2] + ngg, i[1] + ngg, i[0] + ngg].copy())
    return np.concatenate(out) if out else np.zeros(0)


def unpack(bv, f, ngg, data):
    nv = bv.shape[0]
    q = 0
    for idx in _points(f):
        for layer in range(ngg + 1):
            i = list(idx)
            i[f["face"]] += layer * f["direction"]
            sl = (slice(None), i[2] + ngg, i[1] + ngg, i[0] + ngg)
            rec = data[q:q + nv]
            bv[sl] = 0.5 * (bv[sl] + rec) if layer == 0 else rec
            q += nv
    return bv
