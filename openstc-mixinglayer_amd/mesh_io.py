"""Readers and writers of the reference's multi-block connectivity files.

``boundary N.txt`` (Mesh::read_boundary, src/Mesh.cpp:790-810): for every
block a line with its face count, then one line per face
``x1 x2 y1 y2 z1 z2 label`` in 0-based node indices.

``inner N.txt`` (Mesh::read_inner_interface, src/Mesh.cpp:812-843): the
inner-face count of every block (whitespace separated), the rest of that line
and one more line skipped, then per block and face two lines
``i1 i2 j1 j2 k1 k2 id`` -- this block's face, then the target's face with the
target block's 1-based id -- in 1-based node indices, the tangential axis that
pairs with the other side's marked axis written as negative indices.

``parallel N.txt`` (Mesh::read_parallel_interface, src/Mesh.cpp:845-876):
block count and total face count, the per-block face counts, two skipped
lines, then nine integers per face.  Faces here are exchanged between MPI
ranks; a context on one GPU holds every block, so a file that lists any is
refused.

N is the rank right-aligned in four characters (``boundary   0.txt``).
The records become the ``boundaries`` / ``inner_faces`` entries of a case
dict (``range`` / ``target_range`` keys) that ``Solver`` and ``FrSolver``
register through ``ostc_add_boundary_range`` / ``ostc_add_inner_face_range``.
"""
from __future__ import annotations

import os

# boundary labels with an implementation here (src/BoundCond.cuh:1488-1678)
FARFIELD, INFLOW, OUTFLOW, PERIODIC = 4, 5, 6, 8


def file_name(kind: str, rank: int = 0) -> str:
    return f"{kind}{rank:>4}.txt"


def _ints(line: str, n: int, what: str):
    parts = line.split()
    if len(parts) < n:
        raise ValueError(f"{what}: expected {n} integers, got {line!r}")
    return [int(v) for v in parts[:n]]


def read_boundary(path: str, n_block: int):
    """[(block, [x1, x2, y1, y2, z1, z2], label)] in file order."""
    with open(path) as f:
        lines = f.read().splitlines()
    out, pos = [], 0
    for blk in range(n_block):
        if pos >= len(lines):
            raise ValueError(f"{path}: missing the face count of block {blk}")
        n = int(lines[pos].split()[0])
        pos += 1
        for _ in range(n):
            v = _ints(lines[pos], 7, path)
            pos += 1
            out.append((blk, v[:6], v[6]))
    return out


def read_inner(path: str, n_block: int):
    """[(block, range[6], target_range[6], target_block)] in file order;
    target_block is 0-based."""
    with open(path) as f:
        text = f.read()
    lines = text.splitlines()
    # the per-block counts are read token by token, possibly over lines
    counts, li, tok = [], 0, []
    while len(counts) < n_block:
        if not tok:
            if li >= len(lines):
                raise ValueError(f"{path}: missing inner-face counts")
            tok = lines[li].split()
            li += 1
            continue
        counts.append(int(tok.pop(0)))
    # the rest of the current line was consumed above; skip one more line
    li += 1
    out = []
    for blk in range(n_block):
        for _ in range(counts[blk]):
            src = _ints(lines[li], 7, path)
            tgt = _ints(lines[li + 1], 7, path)
            li += 2
            out.append((blk, src[:6], tgt[:6], tgt[6] - 1))
    return out


def read_parallel(path: str, n_block: int):
    """[(block, range[6], rank, send_flag, recv_flag)]."""
    with open(path) as f:
        lines = f.read().splitlines()
    head = lines[0].split()
    if int(head[0]) != n_block:
        raise ValueError(f"{path}: n_block in the file is {head[0]}, the grid has {n_block}")
    tok = head[2:]
    li = 1
    counts = []
    while len(counts) < n_block:
        if not tok:
            tok = lines[li].split()
            li += 1
            continue
        counts.append(int(tok.pop(0)))
    li += 1  # the rest of the counts line, then the header
    rest = " ".join(lines[li:]).split()
    out, q = [], 0
    for blk in range(n_block):
        for _ in range(counts[blk]):
            v = [int(x) for x in rest[q:q + 9]]
            q += 9
            out.append((blk, v[:6], v[6], v[7], v[8]))
    return out


def write_boundary(path: str, n_block: int, records):
    with open(path, "w") as f:
        for blk in range(n_block):
            mine = [r for r in records if r[0] == blk]
            f.write(f"{len(mine):7d}\n")
            for _, rng, label in mine:
                f.write("".join(f"{v:7d}" for v in (*rng, label)) + "\n")


def write_inner(path: str, n_block: int, records):
    with open(path, "w") as f:
        f.write("".join(f"{sum(1 for r in records if r[0] == b):9d}" for b in range(n_block)) + "\n")
        f.write("iMin  iMax  jMin  jMax  kMin  kMax   id\n")
        for blk in range(n_block):
            for _, src, tgt, target in (r for r in records if r[0] == blk):
                f.write("".join(f"{v:7d}" for v in (*src, blk + 1)) + "\n")
                f.write("".join(f"{v:7d}" for v in (*tgt, target + 1)) + "\n")


def write_parallel(path: str, n_block: int, records=()):
    """records: (block, range[6], rank, send_flag, recv_flag), as read_parallel returns them."""
    records = list(records)
    with open(path, "w") as f:
        f.write(f"{n_block:9d}{len(records):9d}\n")
        f.write("".join(f"{sum(1 for r in records if r[0] == b):9d}" for b in range(n_block)) + "\n")
        f.write("iMin  iMax  jMin  jMax  kMin  kMax   s_r_id  send_flag  recv_flag\n")
        for blk in range(n_block):
            for _, rng, rank, fs, fr in (r for r in records if r[0] == blk):
                f.write("".join(f"{v:7d}" for v in (*rng, rank, fs, fr)) + "\n")


def parallel_faces(directory: str, n_block: int, rank: int = 0):
    """The faces shared with other ranks from ``parallel N.txt``, as the
    ``parallel_faces`` records that ``Solver`` and ``mixture.FrSolver``
    register (ostc_add_parallel_face): block, range as in the file, rank, tags."""
    path = os.path.join(directory, file_name("parallel", rank))
    if not os.path.exists(path):
        return []
    return [dict(block=blk, range=list(rng), rank=r, flag_send=fs, flag_recv=fr)
            for blk, rng, r, fs, fr in read_parallel(path, n_block)]


def connectivity(directory: str, n_block: int, rank: int = 0, labels: dict | None = None,
                 farfields: dict | None = None, inflows: dict | None = None, allow_parallel: bool = False):
    """(boundaries, inner_faces) of a case dict from the three files of one
    rank.  ``labels`` maps a boundary label to its type (default: 4 farfield,
    5 inflow, 6 outflow, 8 periodic, the labels of
    src/BoundCond.cuh:1488-1678); ``farfields`` maps a farfield label to its
    state inputs (``farfield_state_air`` / ``FrSolver`` keys), ``inflows`` an
    inflow label to its ``ostc_inflow`` record (``inflow_struct`` keys)."""
    labels = {FARFIELD: FARFIELD, INFLOW: INFLOW, OUTFLOW: OUTFLOW, PERIODIC: PERIODIC} if labels is None else labels
    farfields = farfields or {}
    inflows = inflows or {}
    bnd = read_boundary(os.path.join(directory, file_name("boundary", rank)), n_block)
    inn = read_inner(os.path.join(directory, file_name("inner", rank)), n_block)
    par_path = os.path.join(directory, file_name("parallel", rank))
    if os.path.exists(par_path) and read_parallel(par_path, n_block) and not allow_parallel:
        raise ValueError(f"{par_path} lists faces to other ranks: read them with parallel_faces() "
                         "(allow_parallel=True) for a context that exchanges with them")
    boundaries = []
    for blk, rng, label in bnd:
        if label not in labels:
            raise ValueError(f"boundary label {label} has no implementation here (known: {sorted(labels)})")
        f = dict(block=blk, range=list(rng), type=labels[label], label=label)
        if f["type"] == FARFIELD:
            if label not in farfields:
                raise ValueError(f"farfield label {label} needs its state")
            f["farfield"] = dict(farfields[label])
        if f["type"] == INFLOW:
            if label not in inflows:
                raise ValueError(f"inflow label {label} needs its state")
            f["inflow"] = dict(inflows[label])
        boundaries.append(f)
    inner = [dict(block=blk, range=list(src), target_range=list(tgt), target=t) for blk, src, tgt, t in inn]
    return boundaries, inner
