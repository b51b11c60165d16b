"""The product's inner-face maps (host code, csrc/ostc_faces.hpp) against the
oracle's step-by-step restatement of InnerFace::register_boundary and the
inner_communication loop (oracle/ostc_oracle_faces.h), point by point, on the
CPU: every record of the rotated two-block case, all 48 relative
orientations (permutation x reversal) of a face pair between two boxes, face
pieces, and malformed records that both must refuse.  Built with g++ from
tests/faces_check.cpp.
"""
import itertools
import os
import subprocess
import tempfile

import pytest

from _util import ROOT, cases

CSRC = os.path.join(ROOT, "openstc-mixinglayer_amd", "csrc")
ORACLE = os.path.join(ROOT, "oracle")


@pytest.fixture(scope="module")
def checker():
    d = tempfile.mkdtemp()
    exe = os.path.join(d, "faces_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", CSRC, "-I", ORACLE, "-o", exe,
                    os.path.join(ROOT, "tests", "faces_check.cpp")], check=True)
    return exe


def _run(exe, lines):
    r = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout.split("\n")[:len(lines)]


def _line(ngg, ext, text, src, tgt):
    return " ".join(str(v) for v in (ngg, *ext, *text, *src, *tgt))


def test_rotated_pair_records(checker):
    single = cases.mixing_layer_case([23], [15], [13])
    pair = cases.rotated_pair(single, 10, split_k=7)
    ext = [(b["mx"], b["my"], b["mz"]) for b in pair["blocks"]]
    lines = [_line(3, ext[f["block"]], ext[f["target"]], f["range"], f["target_range"]) for f in pair["inner_faces"]]
    out = _run(checker, lines)
    assert all(o.startswith("ok") for o in out), out


def _record(axis_map, flips, ext):
    """Records of a face pair: block A's high face on axis 0, glued to block
    B whose axis axis_map[d] runs along A's axis d (reversed where flips[d]).
    Returns (extB, src, tgt)."""
    extb = [0, 0, 0]
    for d in range(3):
        extb[axis_map[d]] = ext[d]
    src = [0] * 6
    tgt = [0] * 6
    src[0] = src[1] = ext[0]
    # B's face: where A's high face meets B along B's axis axis_map[0]
    b0 = axis_map[0]
    tface = 1 if not flips[0] else extb[b0]  # B continues A's axis 0 (ascending) or against it
    tgt[2 * b0] = tgt[2 * b0 + 1] = tface
    for d, mark in ((1, -1), (2, 1)):
        bd = axis_map[d]
        src[2 * d], src[2 * d + 1] = mark * 1, mark * ext[d]
        a, b = (1, extb[bd]) if not flips[d] else (extb[bd], 1)
        tgt[2 * bd], tgt[2 * bd + 1] = mark * a, mark * b
    return extb, src, tgt


def test_every_orientation(checker):
    ext = (9, 7, 6)
    lines = []
    for perm in itertools.permutations(range(3)):
        for flips in itertools.product((False, True), repeat=3):
            extb, src, tgt = _record(perm, flips, ext)
            lines.append(_line(3, ext, extb, src, tgt))
            # and a piece of the face: tangential ranges shortened on both sides
            s2, t2 = list(src), list(tgt)
            s2[5] = 4  # z range 1..4 of A
            bd = perm[2]
            a, b = (1, 4) if not flips[2] else (extb[bd], extb[bd] - 3)
            t2[2 * bd], t2[2 * bd + 1] = a, b
            lines.append(_line(3, ext, extb, s2, t2))
    out = _run(checker, lines)
    assert all(o.startswith("ok") for o in out), [(l, o) for l, o in zip(lines, out) if not o.startswith("ok")]
    assert len(out) == 96


def test_malformed_records_are_refused_by_both(checker):
    ext, text = (9, 7, 6), (9, 7, 6)
    good_src, good_tgt = [9, 9, -1, -7, 1, 6], [1, 1, -1, -7, 1, 6]
    lines = [
        _line(3, ext, text, good_src, good_tgt),
        _line(3, ext, text, [9, 9, 1, 7, 1, 6], [1, 1, 1, 7, 1, 6]),      # no marked axis
        _line(3, ext, text, [9, 9, -1, -7, 1, 6], [1, 1, -8, -2, 1, 6]),  # target range leaves its block
        _line(3, ext, (2, 7, 6), good_src, [1, 1, -1, -7, 1, 6]),        # target thinner than ngg+1
        _line(3, ext, text, [9, 9, -1, -7, 1, 6], [1, 1, -1, -6, 1, 6]),  # paired ranges differ in length
    ]
    out = _run(checker, lines)
    assert out[0].startswith("ok") and out[1:4] == ["reject"] * 3, out
    assert out[4].startswith("strict"), out
