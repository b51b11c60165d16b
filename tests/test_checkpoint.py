"""The checkpoint format (flowfield.plt, #!TDV112, as FieldIO writes it,
src/FieldIO.h:72-647) on the CPU: the header fields in the reference's order
read back with struct, the file size equals the header plus
16 + 20 n_var + 8 n_var N bytes per zone (FieldIO::compute_offset_header),
the min/max pairs bound the data, and a write -> read round trip is exact."""
import os
import struct
import tempfile

import numpy as np
import pytest

from _util import pkg

ck = pkg.checkpoint


def _zones(rng, shapes, n_var):
    return [[rng.standard_normal(s) for _ in range(n_var)] for s in shapes]


def test_header_layout_size_and_round_trip():
    rng = np.random.default_rng(5)
    names = ck.BASE_NAMES + ["H2", "O2"]
    shapes = [(4, 5, 6), (3, 7, 2)]
    zones = _zones(rng, shapes, len(names))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "flowfield.plt")
        ck.write_plt(path, names, zones, solution_time=1.25e-3)
        data = open(path, "rb").read()
        back = ck.read_plt(path)
    assert data[:8] == b"#!TDV112"
    assert struct.unpack_from("<ii", data, 8) == (1, 0)
    p = 16
    title = []
    while True:
        c = struct.unpack_from("<i", data, p)[0]
        p += 4
        if c == 0:
            break
        title.append(chr(c))
    assert "".join(title) == "Solution file"
    assert struct.unpack_from("<i", data, p)[0] == len(names)
    # header: magic 8, two ints, title, n_var, names, per zone, EOH
    def slen(s):
        return 4 * (len(s) + 1)
    header = 8 + 8 + slen("Solution file") + 4 + sum(slen(n) for n in names)
    header += sum(4 + slen(f"zone {z}") + 8 + 8 + 20 + 12 + 4 for z in range(len(shapes))) + 4
    assert struct.unpack_from("<f", data, header - 4)[0] == 357.0
    n_var = len(names)
    body = sum(16 + 20 * n_var + 8 * n_var * int(np.prod(s)) for s in shapes)
    assert len(data) == header + body
    # zone 0 data section: marker, formats, min/max of x
    assert struct.unpack_from("<f", data, header)[0] == 299.0
    assert struct.unpack_from(f"<{n_var}i", data, header + 4) == (2,) * n_var
    q = header + 4 + 4 * n_var + 12
    mn, mxv = struct.unpack_from("<dd", data, q)
    assert mn == zones[0][0].min() and mxv == zones[0][0].max()
    assert back["names"] == names and back["time"] == 1.25e-3
    for z, arrays in enumerate(zones):
        assert back["zones"][z]["shape"] == shapes[z]
        for nm, a in zip(names, arrays):
            np.testing.assert_array_equal(back["zones"][z][nm], a)


def test_residual_scale_round_trip():
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "residual_scale.txt")
        ck.write_residual_scale(path, [1.5, 2e-3, 101325.0, 300.0])
        assert ck.read_residual_scale(path) == [1.5, 2e-3, 101325.0, 300.0]


@pytest.mark.parametrize("cut", ["header", "zone_header", "data", "last_byte"])
def test_truncated_file_gives_a_reader_error(cut):
    """A flowfield.plt cut short anywhere -- in the header, in a zone's data
    header, inside the data, or by its last byte -- is refused with a clean
    ValueError naming the truncation, not a struct or numpy error."""
    rng = np.random.default_rng(9)
    names = ck.BASE_NAMES
    zones = _zones(rng, [(3, 4, 5), (2, 3, 4)], len(names))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "flowfield.plt")
        ck.write_plt(path, names, zones)
        data = open(path, "rb").read()
        header_end = data.index(struct.pack("<f", ck.EOH_MARKER)) + 4
        at = {"header": 40, "zone_header": header_end + 10, "data": header_end + 16 + 20 * len(names) + 100,
              "last_byte": len(data) - 1}[cut]
        with open(path, "wb") as f:
            f.write(data[:at])
        with pytest.raises(ValueError, match="truncated"):
            ck.read_plt(path)
