"""Finite-rate mixture path (the reference's MixtureModel::FR) over the C ABI.

* ``load_mechanism(mech, therm, tran, reaction)`` -- Species / Reaction
  constructors, src/ChemData.cpp:11-948 (the library's C++ reader,
  ``ostc_mech_load``);
* ``FrSolver(case, tables)`` -- RK3<MixtureModel::FR> and its operators,
  with the same method names as ``Solver`` (``rk3``, ``compute_rhs``,
  ``global_time_step``, ``update_cv_and_bv_rk``,
  ``update_physical_properties``).

Every call raises ``OstcError`` with the library's message on failure;
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import importlib
import os

import numpy as np

MAX_SPEC, MAX_REAC, MAX_ELEM = 16, 32, 8
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "chemistry")
H2_AIR_FILES = tuple(os.path.join(DATA, f) for f in ("H2_mech.inp", "therm.dat", "tran.dat"))

(FIELD_CV, FIELD_BV, FIELD_SV, FIELD_QN, FIELD_DQ, FIELD_MUL, FIELD_LAMBDA, FIELD_CP, FIELD_GAMMA, FIELD_C,
 FIELD_MACH, FIELD_RHOD, FIELD_JAC, FIELD_METRIC, FIELD_DT_LOCAL) = range(15)
MODE_DIRECT, MODE_GRAPH, MODE_TIMED = 0, 1, 2
# one launch per kernel kind over every block, graph-replayed / direct with per-kind HIP events
MODE_BATCHED, MODE_BATCHED_TIMED = 3, 4
KERNELS = ("source", "inviscid_x", "inviscid_y", "inviscid_z", "viscous_x", "viscous_y", "viscous_z", "update",
           "props")


class MechTables(C.Structure):
    """ostc_mech_tables (include/openstc/mech.h)."""
    S, R, E = MAX_SPEC, MAX_REAC, MAX_ELEM
    _fields_ = [
        ("n_spec", C.c_int), ("n_reac", C.c_int), ("n_elem", C.c_int),
        ("spec_name", (C.c_char * 16) * S), ("elem_name", (C.c_char * 4) * E),
        ("elem_comp", (C.c_int * E) * S), ("mw", C.c_double * S),
        ("t_low", C.c_double * S), ("t_mid", C.c_double * S), ("t_high", C.c_double * S),
        ("low_coeff", (C.c_double * 7) * S), ("high_coeff", (C.c_double * 7) * S),
        ("lj_potent_inv", C.c_double * S), ("vis_coeff", C.c_double * S),
        ("wj_div_wi_to_one4th", (C.c_double * S) * S), ("sqrt_wi_div_wj_pl1_mul8", (C.c_double * S) * S),
        ("binary_diffusivity_coeff", (C.c_double * S) * S), ("kb_over_eps_jk", (C.c_double * S) * S),
        ("label", C.c_int * R), ("rev_type", C.c_int * R), ("order", C.c_int * R),
        ("stoi_f", (C.c_int * S) * R), ("stoi_b", (C.c_int * S) * R),
        ("A", C.c_double * R), ("b", C.c_double * R), ("Ea", C.c_double * R),
        ("A2", C.c_double * R), ("b2", C.c_double * R), ("Ea2", C.c_double * R),
        ("third_body_coeff", (C.c_double * S) * R),
        ("troe_alpha", C.c_double * R), ("troe_t3", C.c_double * R), ("troe_t1", C.c_double * R),
        ("troe_t2", C.c_double * R),
    ]

    def species(self):
        return [self.spec_name[i].value.decode() for i in range(self.n_spec)]


class FrSetup(C.Structure):
    """ostc_fr_setup (include/openstc/abi.h)."""
    _fields_ = [("inviscid_scheme", C.c_int), ("viscous_order", C.c_int), ("reaction", C.c_int),
                ("positive_preserving", C.c_int), ("grad_p_diffusion", C.c_int), ("fixed_time_step", C.c_int),
                ("reconstruction", C.c_int), ("dt", C.c_double), ("cfl", C.c_double), ("prandtl", C.c_double),
                ("rho_ref", C.c_double), ("v_ref", C.c_double), ("entropy_fix_factor", C.c_double),
                ("ch_stencil", C.c_int)]


class _BlockDesc(C.Structure):
    """ostc_block_desc (include/openstc/abi.h), as in the package's mirror."""
    _fields_ = [("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int),
                ("cv", C.c_void_p), ("bv", C.c_void_p), ("sv", C.c_void_p), ("qn", C.c_void_p),
                ("dq", C.c_void_p), ("mul", C.c_void_p), ("jac", C.c_void_p), ("metric", C.c_void_p),
                ("metric_elem_stride", C.c_longlong), ("metric_comp_stride", C.c_longlong),
                ("metric_offset", C.c_longlong)]


_I, _D, _P, _LL = C.c_int, C.c_double, C.c_void_p, C.c_longlong
_DP = C.POINTER(C.c_double)
def _pkg():
    return importlib.import_module(__name__.rsplit(".", 1)[0])


SIGNATURES = {
    "ostc_mech_load": (_I, [C.c_char_p, C.c_char_p, C.c_char_p, _I, C.POINTER(MechTables), C.c_char_p, _I]),
    "ostc_fr_supported": (_I, [_I]),
    "ostc_fr_create": (_I, [C.POINTER(FrSetup), C.POINTER(MechTables), _P, C.POINTER(_P)]),
    "ostc_fr_destroy": (_I, [_P]),
    "ostc_fr_add_block": (_I, [_P, C.POINTER(_BlockDesc), C.POINTER(_I)]),
    "ostc_fr_add_boundary": (_I, [_P, _I, _I, _I, _I]),
    "ostc_fr_add_inner_face": (_I, [_P, _I, _I, _I, _I]),
    "ostc_fr_add_farfield": (_I, [_P, _I, _I, _I, C.POINTER(_pkg().Farfield)]),
    "ostc_fr_add_boundary_range": (_I, [_P, _I, C.POINTER(_I), _I, C.POINTER(_pkg().Farfield)]),
    "ostc_fr_add_inner_face_range": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I), _I]),
    "ostc_fr_add_inflow": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_pkg().Inflow), _DP]),
    "ostc_fr_set_sponge": (_I, [_P, _I, C.c_void_p, _DP, _DP, _DP]),
    "ostc_fr_ngg": (_I, [_P, C.POINTER(_I)]),
    "ostc_fr_store_last_step": (_I, [_P]),
    "ostc_fr_residual": (_I, [_P, _DP]),
    "ostc_fr_farfield_state": (_I, [C.POINTER(MechTables), C.POINTER(_pkg().FarfieldSpec), _DP,
                                    C.POINTER(_pkg().Farfield)]),
    "ostc_fr_finalize": (_I, [_P]),
    "ostc_fr_set_geometry": (_I, [_P, _I, _DP, _DP, _LL, _LL, _LL]),
    "ostc_fr_set_field": (_I, [_P, _I, _I, _DP]),
    "ostc_fr_get_field": (_I, [_P, _I, _I, _DP]),
    "ostc_fr_initialize": (_I, [_P]),
    "ostc_fr_step": (_I, [_P, _I, _I, _DP]),
    "ostc_fr_stage_rhs": (_I, [_P, _I, _D]),
    "ostc_fr_compute_dt": (_I, [_P, _DP]),
    "ostc_fr_stage_update": (_I, [_P, _I, _I, _D]),
    "ostc_fr_update_properties": (_I, [_P, _I]),
    "ostc_fr_copy_qn": (_I, [_P, _I]),
    "ostc_fr_apply_bc": (_I, [_P, _I]),
    "ostc_fr_exchange": (_I, [_P]),
    "ostc_fr_add_parallel_face": (_I, [_P, _I, C.POINTER(_I), _I, _I, _I]),
    "ostc_fr_parallel_face_count": (_I, [_P, C.POINTER(_I)]),
    "ostc_fr_parallel_face_info": (_I, [_P, _I, C.POINTER(_pkg().ParallelInfo)]),
    "ostc_fr_parallel_buffers": (_I, [_P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(_LL)]),
    "ostc_fr_pack_parallel": (_I, [_P]),
    "ostc_fr_unpack_parallel": (_I, [_P]),
    "ostc_fr_parallel_copy": (_I, [_P, _I, _DP]),
    "ostc_fr_last_step_timing": (_I, [_P, _DP, C.POINTER(_LL)]),
    "ostc_fr_kernel_ms": (_I, [_P, _DP, C.POINTER(_LL)]),
    "ostc_fr_sync": (_I, [_P]),
    "ostc_fr_last_error": (C.c_char_p, [_P]),
    "ostc_util_mixing_layer_sv": (_I, [_I, _I, _I, _I, _I, _D, _D, _DP, _DP, _DP]),
}


def _dp(a):
    return a.ctypes.data_as(_DP)


def load_mechanism(mech_file=H2_AIR_FILES[0], therm_file=H2_AIR_FILES[1], tran_file=H2_AIR_FILES[2],
                   reaction=1) -> MechTables:
    """Reads the CHEMKIN mechanism, thermodynamic and transport files with
    the library's reader (ostc_mech_load)."""
    pkg = _pkg()
    t = MechTables()
    err = C.create_string_buffer(512)
    r = pkg.lib().ostc_mech_load(mech_file.encode(), therm_file.encode(), tran_file.encode(), int(reaction),
                                 C.byref(t), err, len(err))
    if r != 0:
        raise pkg.OstcError(f"ostc_mech_load: {err.value.decode()}")
    return t


def supported(n_spec: int) -> bool:
    return bool(_pkg().lib().ostc_fr_supported(n_spec))


def farfield_state(tables: MechTables, y, **kw):
    """FarField constructor with species (src/BoundCond.cpp:586-700): mass
    fractions ``y`` and keyword inputs (mach, pressure, temperature, ...)."""
    pkg = _pkg()
    yy = np.zeros(MAX_SPEC)
    yy[:tables.n_spec] = y
    out = pkg.Farfield()
    r = pkg.lib().ostc_fr_farfield_state(C.byref(tables), C.byref(pkg.farfield_spec(**kw)), _dp(yy), C.byref(out))
    if r != 0:
        raise pkg.OstcError(f"bad far-field specification ({r})")
    return out


class Sponge(C.Structure):
    """ostc_sponge: the sponge layer's sides, ranges and strengths (src/SpongeLayer.cu:28-172)."""
    _fields_ = [("enabled", C.c_int * 6), ("start", C.c_double * 6), ("end", C.c_double * 6),
                ("sigma", C.c_double * 6), ("function", C.c_int), ("jacobian_scaled", C.c_int)]


def sponge_struct(d: dict) -> Sponge:
    return Sponge((C.c_int * 6)(*[int(v) for v in d["enabled"]]), (C.c_double * 6)(*d["start"]),
                  (C.c_double * 6)(*d["end"]), (C.c_double * 6)(*d["sigma"]), int(d.get("function", 0)),
                  int(d.get("jacobian_scaled", 1)))


class FrSolver:
    """Blocks of a finite-rate mixture on one GPU (the reference's layout and
    operators, MixtureModel::FR)."""

    def __init__(self, case: dict, tables: MechTables | None = None, stream: int | None = None,
                 device_arrays=None):
        """device_arrays: per block, a dict of caller-owned device arrays (cv,
        bv, sv, qn, dq, mul, jac, metric + metric_strides; see the package's
        block_desc) that the library adopts instead of allocating its own."""
        pkg = _pkg()
        self.L = L = pkg.lib()
        self.case = case
        self.T = tables if tables is not None else load_mechanism(*case["mech_files"], reaction=case["reaction"])
        self.ns = self.T.n_spec
        self.nv = 5 + self.ns
        s = FrSetup(case["inviscid_scheme"], case["viscous_order"], case["reaction"],
                    int(case.get("positive_preserving", 0)), int(case.get("grad_p_diffusion", 0)),
                    int(case.get("fixed_time_step", 0)), int(case.get("reconstruction", 2)),
                    float(case.get("dt", 0.0)), case["cfl"], case["prandtl"],
                    case["rho_ref"], case["v_ref"], float(case.get("entropy_fix_factor", 0.125)),
                    int(case.get("ch_stencil", 0)))
        ctx = C.c_void_p()
        r = L.ostc_fr_create(C.byref(s), C.byref(self.T), stream, C.byref(ctx))
        if r != 0:
            raise pkg.OstcError(f"ostc_fr_create failed ({r})")
        self.ctx = ctx
        g = C.c_int()
        self._chk(L.ostc_fr_ngg(self.ctx, C.byref(g)))
        self.ngg = g.value
        for i, b in enumerate(case["blocks"]):
            d = _BlockDesc(b["mx"], b["my"], b["mz"])
            for k, v in ((device_arrays[i] if device_arrays else None) or {}).items():
                if k == "metric_strides":
                    d.metric_elem_stride, d.metric_comp_stride, d.metric_offset = v
                else:
                    setattr(d, k, v)
            self._chk(L.ostc_fr_add_block(self.ctx, C.byref(d), None))
        ints = _pkg().ints
        for f in case["boundaries"]:
            if f["type"] == 5:  # inflow: its state and the y of the block's nodes
                blk = case["blocks"][f["block"]]
                y = np.ascontiguousarray(pkg.cases.node_coordinates(case, f["block"], self.ngg)[1])
                self._chk(L.ostc_fr_add_inflow(self.ctx, f["block"], ints(pkg.cases.face_range(blk, f)),
                                               C.byref(pkg.inflow_struct(f["inflow"])),
                                               y.ctypes.data_as(C.POINTER(C.c_double))))
                continue
            ff = farfield_state(self.T, f["farfield_y"], **f["farfield"]) if f["type"] == 4 else None
            if "range" in f:  # a boundary N.txt record
                self._chk(L.ostc_fr_add_boundary_range(self.ctx, f["block"], ints(f["range"]), f["type"],
                                                       None if ff is None else C.byref(ff)))
            elif ff is not None:
                self._chk(L.ostc_fr_add_farfield(self.ctx, f["block"], f["axis"], f["side"], C.byref(ff)))
            else:
                self._chk(L.ostc_fr_add_boundary(self.ctx, f["block"], f["axis"], f["side"], f["type"]))
        for f in case["inner_faces"]:
            if "range" in f:  # an inner N.txt line pair
                self._chk(L.ostc_fr_add_inner_face_range(self.ctx, f["block"], ints(f["range"]),
                                                         ints(f["target_range"]), f["target"]))
            else:
                self._chk(L.ostc_fr_add_inner_face(self.ctx, f["block"], f["axis"], f["side"], f["target"]))
        if "sponge" in case:  # the sponge layer of every block, from its node coordinates
            sp = sponge_struct(case["sponge"])
            for b in range(len(case["blocks"])):
                xyz = [np.ascontiguousarray(a) for a in pkg.cases.node_coordinates(case, b, self.ngg)]
                self._chk(L.ostc_fr_set_sponge(self.ctx, b, C.cast(C.byref(sp), C.c_void_p),
                                               *(a.ctypes.data_as(C.POINTER(C.c_double)) for a in xyz)))
        for f in case.get("parallel_faces", []):  # parallel N.txt records
            self._chk(L.ostc_fr_add_parallel_face(self.ctx, f["block"], ints(f["range"]), f["rank"],
                                                  f.get("flag_send", 0), f.get("flag_recv", 0)))
        self._chk(L.ostc_fr_finalize(self.ctx))

    # -- faces shared with other ranks ------------------------------------
    def parallel_faces(self):
        """[{block, rank, flag_send, flag_recv, offset, count, points}] of the parallel faces."""
        n = C.c_int()
        self._chk(self.L.ostc_fr_parallel_face_count(self.ctx, C.byref(n)))
        out = []
        for q in range(n.value):
            info = _pkg().ParallelInfo()
            self._chk(self.L.ostc_fr_parallel_face_info(self.ctx, q, C.byref(info)))
            out.append({k: getattr(info, k) for k, _ in _pkg().ParallelInfo._fields_})
        return out

    def _parallel_total(self):
        t = _LL()
        self._chk(self.L.ostc_fr_parallel_buffers(self.ctx, None, None, C.byref(t)))
        return t.value

    def pack_parallel(self) -> np.ndarray:
        """Packs every parallel face and returns the send buffer (host copy)."""
        self._chk(self.L.ostc_fr_pack_parallel(self.ctx))
        out = np.zeros(self._parallel_total())
        self._chk(self.L.ostc_fr_parallel_copy(self.ctx, 0, out.ctypes.data_as(_DP)))
        return out

    def unpack_parallel(self, recv: np.ndarray):
        """Copies `recv` into the receive buffer and unpacks every parallel face."""
        recv = np.ascontiguousarray(recv, dtype=np.float64)
        if recv.size != self._parallel_total():
            raise ValueError(f"receive buffer of {recv.size} doubles, expected {self._parallel_total()}")
        self._chk(self.L.ostc_fr_parallel_copy(self.ctx, 1, recv.ctypes.data_as(_DP)))
        self._chk(self.L.ostc_fr_unpack_parallel(self.ctx))

    def _chk(self, r):
        if r != 0:
            raise _pkg().OstcError(f"error {r}: {self.L.ostc_fr_last_error(self.ctx).decode()}")

    def close(self):
        if getattr(self, "ctx", None):
            self.L.ostc_fr_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- fields ------------------------------------------------------------
    def shape(self, b: int, field: int):
        blk = self.case["blocks"][b]
        g = self.ngg
        if field in (FIELD_DQ, FIELD_DT_LOCAL):
            per = (blk["mz"], blk["my"], blk["mx"])
        else:
            per = (blk["mz"] + 2 * g, blk["my"] + 2 * g, blk["mx"] + 2 * g)
        planes = {FIELD_CV: self.nv, FIELD_BV: 6, FIELD_SV: self.ns, FIELD_QN: self.nv, FIELD_DQ: self.nv,
                  FIELD_RHOD: self.ns, FIELD_METRIC: 9}.get(field, 1)
        return (planes, *per)

    def get(self, b: int, field: int) -> np.ndarray:
        out = np.empty(self.shape(b, field))
        self._chk(self.L.ostc_fr_get_field(self.ctx, b, field, _dp(out)))
        return out

    def set(self, b: int, field: int, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.size != int(np.prod(self.shape(b, field))):
            raise _pkg().OstcError("size mismatch")
        self._chk(self.L.ostc_fr_set_field(self.ctx, b, field, _dp(a)))

    def set_geometry(self, b: int, jac: np.ndarray, metric: np.ndarray):
        jac = np.ascontiguousarray(jac, dtype=np.float64)
        metric = np.ascontiguousarray(metric, dtype=np.float64)
        self._chk(self.L.ostc_fr_set_geometry(self.ctx, b, _dp(jac), _dp(metric), 1, jac.size, 0))

    def setup_mixing_layer(self):
        """Uniform-grid geometry, the seeded tanh velocity profile on the two
        stream states and the species of each stream (cases.reacting_case)."""
        pkg = _pkg()
        L, c, g = self.L, self.case, self.ngg
        up = np.array(c["streams"]["upper"], dtype=np.float64)
        lo = np.array(c["streams"]["lower"], dtype=np.float64)
        yu = np.array(c["y_upper"], dtype=np.float64)
        yl = np.array(c["y_lower"], dtype=np.float64)
        per = c["periods"]
        for b, blk in enumerate(c["blocks"]):
            shp = (blk["mz"] + 2 * g, blk["my"] + 2 * g, blk["mx"] + 2 * g)
            jac = np.empty(shp)
            met = np.empty((9, *shp))
            self._chk(L.ostc_util_uniform_geometry(blk["mx"], blk["my"], blk["mz"], g, blk["x0"], blk["y0"],
                                                   blk["z0"], c["dx"], c["dy"], c["dz"], _dp(jac), _dp(met)))
            self.set_geometry(b, jac, met)
            bv = np.empty((6, *shp))
            self._chk(L.ostc_util_mixing_layer_bv(blk["mx"], blk["my"], blk["mz"], g, blk["y0"], c["dy"], _dp(up),
                                                  _dp(lo), c["delta_omega"], blk["gi0"], blk["gj0"], blk["gk0"],
                                                  per[0], per[1], per[2], c["seed"], c["amplitude"], _dp(bv)))
            self.set(b, FIELD_BV, bv)
            sv = np.empty((self.ns, *shp))
            self._chk(L.ostc_util_mixing_layer_sv(blk["mx"], blk["my"], blk["mz"], g, self.ns, blk["y0"], c["dy"],
                                                  _dp(yu), _dp(yl), _dp(sv)))
            self.set(b, FIELD_SV, sv)
        del pkg

    def initialize(self):
        self._chk(self.L.ostc_fr_initialize(self.ctx))

    # -- the reference's operators ---------------------------------------
    def rk3(self, n_steps: int = 1, mode: int = MODE_GRAPH, want_dt: bool = True) -> float:
        dt = C.c_double(0.0)
        self._chk(self.L.ostc_fr_step(self.ctx, n_steps, mode, C.byref(dt) if want_dt else None))
        return dt.value

    def compute_rhs(self, b: int, dt: float = 0.0):
        """dq = J omega; inviscid x, y, z; viscous x, y, z (dt feeds the
        positivity limiter)."""
        self._chk(self.L.ostc_fr_stage_rhs(self.ctx, b, dt))

    def global_time_step(self) -> float:
        dt = C.c_double()
        self._chk(self.L.ostc_fr_compute_dt(self.ctx, C.byref(dt)))
        return dt.value

    def update_cv_and_bv_rk(self, b: int, rk: int, dt: float):
        self._chk(self.L.ostc_fr_stage_update(self.ctx, b, rk, dt))

    def update_physical_properties(self, b: int):
        self._chk(self.L.ostc_fr_update_properties(self.ctx, b))

    def store_last_step(self):
        """store_last_step, src/TimeAdvanceFunc.cu:6-21."""
        self._chk(self.L.ostc_fr_store_last_step(self.ctx))

    def compute_residual(self) -> np.ndarray:
        """Unscaled RMS change of rho, |V|, p, T since store_last_step
        (src/Residual.cuh:12-187)."""
        r = np.zeros(4)
        self._chk(self.L.ostc_fr_residual(self.ctx, _dp(r)))
        return r

    def last_step_timing(self):
        ms, n = C.c_double(), C.c_longlong()
        self._chk(self.L.ostc_fr_last_step_timing(self.ctx, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def kernel_ms(self):
        """{kernel kind: (summed ms, launches)} of the last MODE_TIMED or MODE_BATCHED_TIMED call."""
        ms = np.zeros(len(KERNELS))
        cnt = (C.c_longlong * len(KERNELS))()
        self._chk(self.L.ostc_fr_kernel_ms(self.ctx, _dp(ms), cnt))
        return {k: (ms[i], cnt[i]) for i, k in enumerate(KERNELS)}

    def sync(self):
        self._chk(self.L.ostc_fr_sync(self.ctx))
