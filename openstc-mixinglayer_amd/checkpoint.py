"""Checkpoints in the reference's on-disk format.

``flowfield.plt`` is a Tecplot binary file, version ``#!TDV112``, written the
way FieldIO writes it (src/FieldIO.h:72-447, print_field :491-647) with the
default output settings (no ghost layers, no user-defined variables):

header
    "#!TDV112" (8 bytes), int32 1, int32 0 (full file), the title
    "Solution file", int32 n_var, the variable names -- x, y, z, density, u,
    v, w, pressure, temperature, mach, then the species names of a mixture;
    per zone: float32 299.0, "zone i", int32 -1 (parent), int32 -2 (strand),
    float64 solution time, int32 -1 (colour), int32 0 (ordered), int32 0
    (nodal), int32 0, int32 0, int32 IMax JMax KMax, int32 0 (no auxiliary
    data); float32 357.0 (end of header).  Strings are one int32 per
    character and an int32 0.
data, per zone
    float32 299.0, int32 2 (double) per variable, int32 0 (no passive
    variables), int32 0 (no sharing), int32 -1 (no connectivity sharing),
    float64 min and max of every variable, then every variable as IMax x JMax
    x KMax doubles, i fastest.  Each zone takes 16 + 20 n_var + 8 n_var N
    bytes (FieldIO::compute_offset_header).

``residual_scale.txt`` holds the four residual scales of compute_residual
(src/Residual.cuh:140-150), read back on restart (src/Driver.cu:33-39).

A restart loads density, u, v, w, pressure, temperature (and the mass
fractions) of every block's nodes -- and the ghost layers the file holds --
and calls initialize(): boundary conditions, exchange, properties and cv
from the primitives, as the reference's restart does
(src/Initialize.cuh:103-352).
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"#!TDV112"
ZONE_MARKER = 299.0
EOH_MARKER = 357.0
BASE_NAMES = ["x", "y", "z", "density", "u", "v", "w", "pressure", "temperature", "mach"]


def _str(s: str) -> bytes:
    return b"".join(struct.pack("<i", ord(c)) for c in s) + struct.pack("<i", 0)


def write_plt(path: str, names, zones, solution_time: float = 0.0):
    """names: variable names; zones: per zone a list of arrays (one per
    name), each shaped (KMax, JMax, IMax)."""
    n_var = len(names)
    out = bytearray()
    out += MAGIC
    out += struct.pack("<ii", 1, 0)
    out += _str("Solution file")
    out += struct.pack("<i", n_var)
    for nm in names:
        out += _str(nm)
    for z, arrays in enumerate(zones):
        kmax, jmax, imax = np.shape(arrays[0])
        out += struct.pack("<f", ZONE_MARKER)
        out += _str(f"zone {z}")
        out += struct.pack("<ii", -1, -2)
        out += struct.pack("<d", solution_time)
        out += struct.pack("<iiiii", -1, 0, 0, 0, 0)
        out += struct.pack("<iii", imax, jmax, kmax)
        out += struct.pack("<i", 0)
    out += struct.pack("<f", EOH_MARKER)
    for arrays in zones:
        if len(arrays) != n_var:
            raise ValueError("every zone needs one array per variable")
        out += struct.pack("<f", ZONE_MARKER)
        out += struct.pack(f"<{n_var}i", *([2] * n_var))
        out += struct.pack("<iii", 0, 0, -1)
        for a in arrays:
            out += struct.pack("<dd", float(np.min(a)), float(np.max(a)))
        for a in arrays:
            out += np.ascontiguousarray(a, dtype="<f8").tobytes()
    with open(path, "wb") as f:
        f.write(out)


class _Reader:
    def __init__(self, data: bytes, path: str = ""):
        self.d, self.p, self.path = data, 0, path

    def need(self, n: int):
        if self.p + n > len(self.d):
            raise ValueError(f"{self.path}: truncated at byte {len(self.d)} (reading {n} bytes at {self.p})")

    def take(self, fmt: str):
        self.need(struct.calcsize(fmt))
        v = struct.unpack_from(fmt, self.d, self.p)
        self.p += struct.calcsize(fmt)
        return v if len(v) > 1 else v[0]

    def string(self) -> str:
        chars = []
        while True:
            c = self.take("<i")
            if c == 0:
                return "".join(chars)
            chars.append(chr(c))


def read_plt(path: str):
    """{'names', 'time', 'zones': [{'shape': (K, J, I), name: array, ...}]}
    of a file in the layout above (ordered zones, double data)."""
    with open(path, "rb") as f:
        r = _Reader(f.read(), path)
    if r.d[:8] != MAGIC:
        raise ValueError(f"{path}: not a #!TDV112 file")
    r.p = 8
    if r.take("<i") != 1:
        raise ValueError(f"{path}: byte order")
    r.take("<i")
    r.string()
    n_var = r.take("<i")
    names = [r.string() for _ in range(n_var)]
    dims, time = [], 0.0
    while True:
        marker = r.take("<f")
        if marker == EOH_MARKER:
            break
        if marker != ZONE_MARKER:
            raise ValueError(f"{path}: bad zone marker {marker}")
        r.string()
        r.take("<ii")
        time = r.take("<d")
        if r.take("<i") != -1 or r.take("<i") != 0:
            raise ValueError(f"{path}: only ordered zones are read")
        r.take("<iii")
        imax, jmax, kmax = r.take("<iii")
        if r.take("<i") != 0:
            raise ValueError(f"{path}: auxiliary zone data is not read")
        dims.append((kmax, jmax, imax))
    zones = []
    for shape in dims:
        if r.take("<f") != ZONE_MARKER:
            raise ValueError(f"{path}: bad data-section zone marker")
        r.need(16 * n_var)
        fmts = r.take(f"<{n_var}i") if n_var > 1 else (r.take("<i"),)
        if any(v != 2 for v in fmts):
            raise ValueError(f"{path}: only double data is read")
        r.take("<iii")
        r.p += 16 * n_var  # min/max pairs
        n = shape[0] * shape[1] * shape[2]
        zone = {"shape": shape}
        for nm in names:
            r.need(8 * n)
            zone[nm] = np.frombuffer(r.d, dtype="<f8", count=n, offset=r.p).reshape(shape).copy()
            r.p += 8 * n
        zones.append(zone)
    return {"names": names, "time": time, "zones": zones}


def write_residual_scale(path: str, scales):
    with open(path, "w") as f:
        f.write(" ".join(repr(float(v)) for v in scales) + "\n")


def read_residual_scale(path: str):
    with open(path) as f:
        return [float(v) for v in f.read().split()[:4]]


def _inner(a, g, keep=0):
    """The nodes of a padded array with `keep` of its g ghost layers."""
    c = g - keep
    return a[..., c:a.shape[-3] - c, c:a.shape[-2] - c, c:a.shape[-1] - c]


def save(solver, path: str, solution_time: float = 0.0, ghosts: int = 0):
    """flowfield.plt of a Solver or mixture FrSolver: node coordinates of the
    case's grid, the primitives, Mach and (for a mixture) the mass fractions
    under the species names of its mechanism, on every block's nodes and
    `ghosts` ghost layers (the reference's ngg_output, 0 by default).  Edge
    and corner ghost nodes take no part in the exchange, so only a file with
    all ngg ghost layers restarts a run where it stopped, to round-off."""
    from . import cases
    mx = getattr(solver, "T", None) is not None
    species = solver.T.species() if mx else []
    g = solver.ngg
    if mx:
        from . import mixture as M
        f_bv, f_mach, f_sv = M.FIELD_BV, M.FIELD_MACH, M.FIELD_SV
    else:
        from . import FIELD_BV as f_bv, FIELD_MACH as f_mach
        f_sv = None
    names = BASE_NAMES + list(species)
    zones = []
    if not 0 <= ghosts <= g:
        raise ValueError(f"ghosts must be in 0..{g}")
    for b in range(len(solver.case["blocks"])):
        xyz = cases.node_coordinates(solver.case, b, ghosts)
        bv = _inner(solver.get(b, f_bv), g, ghosts)
        mach = _inner(solver.get(b, f_mach), g, ghosts)[0]
        arrays = list(xyz) + [bv[l] for l in range(6)] + [mach]
        if species:
            sv = _inner(solver.get(b, f_sv), g, ghosts)
            arrays += [sv[l] for l in range(len(species))]
        zones.append(arrays)
    write_plt(path, names, zones, solution_time)


def load(solver, path: str):
    """Restart a Solver or FrSolver from flowfield.plt: the primitives (and
    mass fractions, by species name) of every block's nodes, then
    initialize().  Returns the solution time."""
    data = read_plt(path)
    mx = getattr(solver, "T", None) is not None
    species = solver.T.species() if mx else []
    g = solver.ngg
    if mx:
        from . import mixture as M
        f_bv, f_sv = M.FIELD_BV, M.FIELD_SV
    else:
        from . import FIELD_BV as f_bv
        f_sv = None
    if len(data["zones"]) != len(solver.case["blocks"]):
        raise ValueError(f"{path}: {len(data['zones'])} zones for {len(solver.case['blocks'])} blocks")
    for b, zone in enumerate(data["zones"]):
        blk = solver.case["blocks"][b]
        keep = (zone["shape"][2] - blk["mx"]) // 2   # ghost layers in the file
        bv = solver.get(b, f_bv)
        inner = _inner(bv, g, keep) if 0 <= keep <= g else None
        if inner is None or inner.shape[1:] != zone["shape"]:
            raise ValueError(f"{path}: zone {b} has shape {zone['shape']}, block {bv.shape[1:]} with {g} ghosts")
        for l, nm in enumerate(BASE_NAMES[3:9]):
            inner[l] = zone[nm]
        solver.set(b, f_bv, bv)
        if species:
            sv = solver.get(b, f_sv)
            si = _inner(sv, g, keep)
            for l, nm in enumerate(species):
                si[l] = zone[nm]
            solver.set(b, f_sv, sv)
    solver.initialize()
    return data["time"]
