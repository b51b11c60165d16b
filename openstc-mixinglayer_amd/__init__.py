"""Host-side mirror of OpenSTC's RHS + SSP-RK3 operator surface over the C ABI
of libostc_hip.so (include/openstc/abi.h).

The reference drives these operators from ``RK3<mix,turb>(Driver&)``
(src/RK.cuh:20-228); this module exposes them under the same names for
Python callers (tests, benchmarks, the adjoint wrapper):

* ``Solver.rk3(n_steps, mode)``                        -- src/RK.cuh:85-226
* ``Solver.compute_rhs(b)``  (inviscid + viscous dq)    -- src/SchemeSelector.cuh:24-67
* ``Solver.global_time_step()``                        -- src/TimeAdvanceFunc.cu:39-64
* ``Solver.update_cv_and_bv_rk(b, rk, dt)``             -- src/RK.cuh:230-284
* ``Solver.apply_boundary_conditions(b)``               -- src/BoundCond.cuh:1488-1678
* ``Solver.data_communication()``                      -- src/DataCommunication.cuh:29-58
* ``Solver.update_physical_properties(b)``              -- src/FieldOperation.cuh:107-140
* ``Solver.compute_residual()``                        -- src/Residual.cuh:75-187
* ``mixture.FrSolver`` -- the same operators for finite-rate mixtures
  (MixtureModel::FR), ``mixture.load_mechanism`` -- src/ChemData.cpp:11-948

Where the reference aborts the run (MPI_Abort after cudaGetLastError), these
raise ``OstcError`` carrying the library's message.  There is no CPU
fallback: without the built library every call raises.
"""
from __future__ import annotations

import ctypes as C
import importlib
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libostc_hip.so")

cases = importlib.import_module(__name__ + ".cases")

MODE_PLAIN, MODE_BATCHED, MODE_BATCHED_TIMED = 0, 1, 2
GHOST_PROPS_IN_OPS, GHOST_PROPS_PASS = 0, 1
(FIELD_CV, FIELD_BV, FIELD_MUL, FIELD_MACH, FIELD_DQ, FIELD_DT_LOCAL, FIELD_JAC, FIELD_METRIC, FIELD_QN,
 FIELD_DQ_BAR, FIELD_CV_BAR) = range(11)
MAX_MODES = 64


class OstcError(RuntimeError):
    pass


class _Setup(C.Structure):
    _fields_ = [("n_var", C.c_int), ("inviscid_scheme", C.c_int), ("viscous_order", C.c_int),
                ("fixed_time_step", C.c_int), ("dt", C.c_double), ("cfl", C.c_double),
                ("prandtl", C.c_double), ("rho_ref", C.c_double), ("v_ref", C.c_double),
                ("reconstruction", C.c_int), ("entropy_fix_factor", C.c_double)]


class Farfield(C.Structure):
    """ostc_farfield: FarField members after its constructor (src/BoundCond.h:129-149)."""
    _fields_ = [("density", C.c_double), ("u", C.c_double), ("v", C.c_double), ("w", C.c_double),
                ("pressure", C.c_double), ("temperature", C.c_double), ("acoustic_speed", C.c_double),
                ("entropy", C.c_double), ("specific_heat_ratio", C.c_double), ("mw", C.c_double),
                ("sv", C.c_double * 16)]


class Inflow(C.Structure):
    """ostc_inflow: Inflow members (src/BoundCond.cpp:10-321) and the white-noise seed."""
    _fields_ = [("inflow_type", C.c_int), ("fluctuation_type", C.c_int), ("delta_omega", C.c_double),
                ("fluctuation_intensity", C.c_double), ("seed", C.c_uint64), ("upper", C.c_double * 6),
                ("lower", C.c_double * 6), ("sv_upper", C.c_double * 16), ("sv_lower", C.c_double * 16)]


def inflow_struct(d: dict) -> Inflow:
    """An ostc_inflow from a case's ``inflow`` record."""
    sv_u = list(d.get("sv_upper", [])) + [0.0] * (16 - len(d.get("sv_upper", [])))
    sv_l = list(d.get("sv_lower", [])) + [0.0] * (16 - len(d.get("sv_lower", [])))
    return Inflow(int(d.get("inflow_type", 2)), int(d.get("fluctuation_type", 0)), float(d.get("delta_omega", 0.0)),
                  float(d.get("fluctuation_intensity", 0.0)), int(d.get("seed", 0)),
                  (C.c_double * 6)(*d["upper"]), (C.c_double * 6)(*d["lower"]),
                  (C.c_double * 16)(*sv_u), (C.c_double * 16)(*sv_l))


class FarfieldSpec(C.Structure):
    """ostc_farfield_spec: FarField constructor inputs (src/BoundCond.cpp:586-700)."""
    _fields_ = [("mach", C.c_double), ("pressure", C.c_double), ("temperature", C.c_double),
                ("velocity", C.c_double), ("density", C.c_double), ("u", C.c_double), ("v", C.c_double),
                ("w", C.c_double), ("reynolds_number", C.c_double)]


def farfield_spec(**kw) -> FarfieldSpec:
    """Constructor inputs with the reference's defaults for keys not given."""
    d = dict(mach=-1.0, pressure=-1.0, temperature=-1.0, velocity=0.0, density=-1.0, u=1.0, v=0.0, w=0.0,
             reynolds_number=-1.0)
    d.update(kw)
    return FarfieldSpec(**{k: float(v) for k, v in d.items()})


def farfield_state_air(**kw) -> Farfield:
    """FarField::FarField for air from keyword inputs (mach, pressure, ...)."""
    out = Farfield()
    r = lib().ostc_farfield_state_air(C.byref(farfield_spec(**kw)), C.byref(out))
    if r != 0:
        raise OstcError(f"bad far-field specification ({r})")
    return out


class _BlockDesc(C.Structure):
    """ostc_block_desc (include/openstc/abi.h): extents and the optional
    caller-owned device arrays (0: the library allocates)."""
    _fields_ = [("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int),
                ("cv", C.c_void_p), ("bv", C.c_void_p), ("sv", C.c_void_p), ("qn", C.c_void_p),
                ("dq", C.c_void_p), ("mul", C.c_void_p), ("jac", C.c_void_p), ("metric", C.c_void_p),
                ("metric_elem_stride", C.c_longlong), ("metric_comp_stride", C.c_longlong),
                ("metric_offset", C.c_longlong)]


def block_desc(blk: dict, arrays: dict | None = None) -> _BlockDesc:
    """The descriptor of a case block; ``arrays`` maps cv, bv, sv, qn, dq,
    mul, jac and metric to device addresses the caller owns, and
    metric_strides to (element, component, offset) of the metric."""
    d = _BlockDesc(blk["mx"], blk["my"], blk["mz"])
    for k, v in (arrays or {}).items():
        if k == "metric_strides":
            d.metric_elem_stride, d.metric_comp_stride, d.metric_offset = v
        else:
            setattr(d, k, v)
    return d


class Zone(C.Structure):
    """Device view of a block (the DZone equivalent)."""
    _fields_ = [("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int), ("ngg", C.c_int), ("sz", C.c_longlong),
                ("cv", C.c_void_p), ("bv", C.c_void_p), ("mul", C.c_void_p), ("mach", C.c_void_p),
                ("jac", C.c_void_p), ("metric", C.c_void_p), ("dq", C.c_void_p), ("dt_local", C.c_void_p)]


class ParallelInfo(C.Structure):
    """ostc_parallel_info: one face shared with another rank."""
    _fields_ = [("block", C.c_int), ("rank", C.c_int), ("flag_send", C.c_int), ("flag_recv", C.c_int),
                ("offset", C.c_longlong), ("count", C.c_longlong), ("points", C.c_longlong)]


class Fourier(C.Structure):
    """ostc_fourier: Fourier-mode velocity perturbation (differentiation workload)."""
    _fields_ = [("n_modes", C.c_int), ("comp", C.c_int * MAX_MODES), ("kx", C.c_int * MAX_MODES),
                ("kz", C.c_int * MAX_MODES), ("trig", C.c_int * MAX_MODES), ("amp", C.c_double * MAX_MODES),
                ("y0", C.c_double), ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("lx", C.c_double), ("lz", C.c_double), ("delta", C.c_double), ("yc", C.c_double),
                ("du", C.c_double)]


class Theta(C.Structure):
    """ostc_theta: momentum-thickness objective."""
    _fields_ = [("u1", C.c_double), ("u2", C.c_double), ("rho0", C.c_double), ("du", C.c_double),
                ("dy", C.c_double), ("ni", C.c_int), ("nk", C.c_int)]


def fourier_struct(modes: dict, amp=None) -> Fourier:
    """ostc_fourier from a case's ``modes`` dict (cases.fourier_modes); ``amp``
    overrides the amplitudes."""
    f = Fourier()
    a = modes["amp"] if amp is None else amp
    f.n_modes = len(a)
    for m in range(len(a)):
        f.comp[m], f.kx[m], f.kz[m], f.trig[m] = (int(modes[k][m]) for k in ("comp", "kx", "kz", "trig"))
        f.amp[m] = float(a[m])
    for k in ("y0", "dx", "dy", "dz", "lx", "lz", "delta", "yc", "du"):
        setattr(f, k, float(modes[k]))
    return f


class DualTime(C.Structure):
    """ostc_dual_time: dual time stepping + DPLUR keys of setup.txt."""
    _fields_ = [("dt", C.c_double), ("inner_iteration", C.c_int), ("max_inner_iteration", C.c_int),
                ("iteration_adjust_step", C.c_int), ("dplur_inner_step", C.c_int), ("step", C.c_int)]


def dual_time_struct(d: dict) -> DualTime:
    """The reference's defaults (src/Parameter.cpp:449-468) for keys not given."""
    return DualTime(d["dt"], d.get("inner_iteration", 20), d.get("max_inner_iteration", 30),
                    d.get("iteration_adjust_step", 20), d.get("dplur_inner_step", 3), d.get("step", 0))


def theta_struct(t: dict) -> Theta:
    return Theta(t["u1"], t["u2"], t["rho0"], t["du"], t["dy"], t["ni"], t["nk"])


_lib = None

# every exported symbol with its (restype, argtypes)
_I, _D, _P, _LL = C.c_int, C.c_double, C.c_void_p, C.c_longlong
_DP = C.POINTER(C.c_double)
SIGNATURES = {
    "ostc_create": (_I, [C.POINTER(_Setup), _P, C.POINTER(_P)]),
    "ostc_destroy": (_I, [_P]),
    "ostc_add_block": (_I, [_P, C.POINTER(_BlockDesc), C.POINTER(_I)]),
    "ostc_add_boundary": (_I, [_P, _I, _I, _I, _I]),
    "ostc_add_farfield": (_I, [_P, _I, _I, _I, C.POINTER(Farfield)]),
    "ostc_farfield_state_air": (_I, [C.POINTER(FarfieldSpec), C.POINTER(Farfield)]),
    "ostc_add_inner_face": (_I, [_P, _I, _I, _I, _I]),
    "ostc_add_boundary_range": (_I, [_P, _I, C.POINTER(_I), _I, C.POINTER(Farfield)]),
    "ostc_add_inner_face_range": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I), _I]),
    "ostc_add_inflow": (_I, [_P, _I, C.POINTER(_I), C.POINTER(Inflow), _DP]),
    "ostc_set_sponge": (_I, [_P, _I, _P, _DP, _DP, _DP]),
    "ostc_finalize": (_I, [_P]),
    "ostc_add_parallel_face": (_I, [_P, _I, C.POINTER(_I), _I, _I, _I]),
    "ostc_parallel_face_count": (_I, [_P, C.POINTER(_I)]),
    "ostc_parallel_face_info": (_I, [_P, _I, C.POINTER(ParallelInfo)]),
    "ostc_parallel_buffers": (_I, [_P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_longlong)]),
    "ostc_pack_parallel": (_I, [_P]),
    "ostc_unpack_parallel": (_I, [_P]),
    "ostc_parallel_copy": (_I, [_P, _I, _DP]),
    "ostc_ngg": (_I, [_P, C.POINTER(_I)]),
    "ostc_zone_view": (_I, [_P, _I, C.POINTER(Zone)]),
    "ostc_set_geometry": (_I, [_P, _I, _DP, _DP, _LL, _LL, _LL]),
    "ostc_set_field": (_I, [_P, _I, _I, _DP]),
    "ostc_get_field": (_I, [_P, _I, _I, _DP]),
    "ostc_initialize": (_I, [_P]),
    "ostc_step": (_I, [_P, _I, _I, _DP]),
    "ostc_stage_rhs": (_I, [_P, _I]),
    "ostc_compute_dt": (_I, [_P, _DP]),
    "ostc_stage_update": (_I, [_P, _I, _I, _D]),
    "ostc_apply_bc": (_I, [_P, _I]),
    "ostc_exchange": (_I, [_P]),
    "ostc_update_properties": (_I, [_P, _I]),
    "ostc_copy_qn": (_I, [_P, _I]),
    "ostc_batched_rhs": (_I, [_P, _I]),
    "ostc_store_last_step": (_I, [_P]),
    "ostc_residual": (_I, [_P, _DP]),
    "ostc_tile_count": (_I, [C.POINTER(_I)]),
    "ostc_tile_shape": (_I, [_I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "ostc_set_tile": (_I, [_P, _I, _I, _I]),
    "ostc_set_ghost_props": (_I, [_P, _I]),
    "ostc_tile_config": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "ostc_autotune": (_I, [_P, _I, C.POINTER(_I), _DP]),
    "ostc_last_step_timing": (_I, [_P, _DP, C.POINTER(_LL)]),
    "ostc_fused_kernel_ms": (_I, [_P, _DP, _DP, C.POINTER(_LL)]),
    "ostc_sync": (_I, [_P]),
    "ostc_last_error": (C.c_char_p, [_P]),
    "ostc_util_uniform_geometry": (_I, [_I, _I, _I, _I, _D, _D, _D, _D, _D, _D, _DP, _DP]),
    "ostc_util_metric": (_I, [_I, _I, _I, _I, _DP, _DP, _DP, _DP, _DP]),
    "ostc_util_mixing_layer_bv": (_I, [_I, _I, _I, _I, _D, _D, _DP, _DP, _D, _I, _I, _I, _I, _I, _I,
                                       C.c_uint64, _D, _DP]),
    "ostc_rhs_vjp": (_I, [_P, _I, _P, _P]),
    "ostc_perturb_fourier": (_I, [_P, _I, C.POINTER(Fourier)]),
    "ostc_momentum_thickness": (_I, [_P, _I, C.POINTER(Theta), _DP]),
    "ostc_set_dual_time": (_I, [_P, C.POINTER(DualTime)]),
    "ostc_dual_time_step": (_I, [_P, _I, C.POINTER(C.c_int)]),
    "ostc_momentum_thickness_noalloc": (_I, [_P, _I, C.POINTER(Theta), _DP]),
    "ostc_adjoint_forward": (_I, [_P, _I, _I, C.POINTER(Theta), _DP]),
    "ostc_adjoint_backward": (_I, [_P, C.POINTER(Fourier), _DP]),
    "ostc_adjoint_timing": (_I, [_P, _DP, _DP]),
}

# the finite-rate mixture path (MixtureModel::FR): FrSolver, load_mechanism
mixture = importlib.import_module(__name__ + ".mixture")
# readers and writers of the reference's connectivity files
mesh_io = importlib.import_module(__name__ + ".mesh_io")
# checkpoints in the reference's Tecplot (#!TDV112) format
checkpoint = importlib.import_module(__name__ + ".checkpoint")
# the reverse mode as a torch.autograd.Function (imports torch on first use)
autodiff = importlib.import_module(__name__ + ".autodiff")
SIGNATURES.update(mixture.SIGNATURES)


def lib():
    """Loads libostc_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OstcError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_DP)


def ints(v):
    """A C int array of a connectivity record."""
    return (C.c_int * len(v))(*[int(x) for x in v])


def metric_from_coordinates(mx: int, my: int, mz: int, g: int, x: np.ndarray, y: np.ndarray, z: np.ndarray):
    """(jac, metric) on the domain padded by g from node coordinates on the
    domain padded by g+1 (Block::compute_jac_metric, src/Mesh.cpp:241-502)."""
    shp = (mz + 2 * g, my + 2 * g, mx + 2 * g)
    xs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z)]
    if any(a.size != (mz + 2 * g + 2) * (my + 2 * g + 2) * (mx + 2 * g + 2) for a in xs):
        raise OstcError("coordinates must cover the domain padded by g+1 layers")
    jac, met = np.empty(shp), np.empty((9, *shp))
    r = lib().ostc_util_metric(mx, my, mz, g, *(_dp(a) for a in xs), _dp(jac), _dp(met))
    if r != 0:
        raise OstcError(f"ostc_util_metric failed ({r}): non-positive Jacobian")
    return jac, met


def padded_shape(blk: dict, g: int):
    return (blk["mz"] + 2 * g, blk["my"] + 2 * g, blk["mx"] + 2 * g)


class Solver:
    """A set of blocks on one GPU with the reference's layout and operators."""

    def __init__(self, case: dict, stream: int | None = None, device_arrays=None):
        """device_arrays: per block, a dict of caller-owned device arrays
        (block_desc) that the library adopts instead of allocating its own."""
        L = lib()
        self.case = case
        s = _Setup(case["n_var"], case["inviscid_scheme"], case["viscous_order"],
                   int(case.get("fixed_time_step", 0)), float(case.get("dt", 0.0)), case["cfl"],
                   case["prandtl"], case["rho_ref"], case["v_ref"], int(case.get("reconstruction", 2)),
                   float(case.get("entropy_fix_factor", 0.125)))
        ctx = C.c_void_p()
        r = L.ostc_create(C.byref(s), stream, C.byref(ctx))
        if r != 0:
            raise OstcError(f"ostc_create failed ({r})")
        self.ctx = ctx
        g = C.c_int()
        L.ostc_ngg(self.ctx, C.byref(g))
        self.ngg = g.value
        for i, b in enumerate(case["blocks"]):
            arrays = device_arrays[i] if device_arrays else None
            self._chk(L.ostc_add_block(self.ctx, C.byref(block_desc(b, arrays)), None))
        for f in case["boundaries"]:
            if f["type"] == 5:  # inflow: its state and the y of the block's nodes
                blk = case["blocks"][f["block"]]
                y = np.ascontiguousarray(cases.node_coordinates(case, f["block"], self.ngg)[1])
                self._chk(L.ostc_add_inflow(self.ctx, f["block"], ints(cases.face_range(blk, f)),
                                            C.byref(inflow_struct(f["inflow"])), _dp(y)))
                continue
            ff = farfield_state_air(**f["farfield"]) if f["type"] == 4 else None
            if "range" in f:  # a boundary N.txt record
                self._chk(L.ostc_add_boundary_range(self.ctx, f["block"], ints(f["range"]), f["type"],
                                                    None if ff is None else C.byref(ff)))
            elif ff is not None:
                self._chk(L.ostc_add_farfield(self.ctx, f["block"], f["axis"], f["side"], C.byref(ff)))
            else:
                self._chk(L.ostc_add_boundary(self.ctx, f["block"], f["axis"], f["side"], f["type"]))
        for f in case["inner_faces"]:
            if "range" in f:  # an inner N.txt line pair
                self._chk(L.ostc_add_inner_face_range(self.ctx, f["block"], ints(f["range"]),
                                                      ints(f["target_range"]), f["target"]))
            else:
                self._chk(L.ostc_add_inner_face(self.ctx, f["block"], f["axis"], f["side"], f["target"]))
        if "sponge" in case:  # the sponge layer of every block, from its node coordinates
            sp = mixture.sponge_struct(case["sponge"])
            for b in range(len(case["blocks"])):
                xyz = [np.ascontiguousarray(a) for a in cases.node_coordinates(case, b, self.ngg)]
                self._chk(L.ostc_set_sponge(self.ctx, b, C.cast(C.byref(sp), C.c_void_p), *(_dp(a) for a in xyz)))
        for f in case.get("parallel_faces", []):  # parallel N.txt records
            self._chk(L.ostc_add_parallel_face(self.ctx, f["block"], ints(f["range"]), f["rank"],
                                               f.get("flag_send", 0), f.get("flag_recv", 0)))
        self._chk(L.ostc_finalize(self.ctx))

    def _chk(self, r):
        if r != 0:
            raise OstcError(f"error {r}: {lib().ostc_last_error(self.ctx).decode()}")

    def close(self):
        if getattr(self, "ctx", None):
            lib().ostc_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- geometry and state ---------------------------------------------
    def shape(self, b: int, field: int):
        blk = self.case["blocks"][b]
        if field in (FIELD_DQ, FIELD_DT_LOCAL, FIELD_DQ_BAR):
            per = (blk["mz"], blk["my"], blk["mx"])
        else:
            per = padded_shape(blk, self.ngg)
        planes = {FIELD_CV: 5, FIELD_BV: 6, FIELD_MUL: 1, FIELD_MACH: 1, FIELD_DQ: 5, FIELD_DT_LOCAL: 1,
                  FIELD_JAC: 1, FIELD_METRIC: 9, FIELD_QN: 5, FIELD_DQ_BAR: 5, FIELD_CV_BAR: 5}[field]
        return (planes, *per)

    def get(self, b: int, field: int) -> np.ndarray:
        out = np.empty(self.shape(b, field))
        self._chk(lib().ostc_get_field(self.ctx, b, field, _dp(out)))
        return out

    def set(self, b: int, field: int, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.size != int(np.prod(self.shape(b, field))):
            raise OstcError("size mismatch")
        self._chk(lib().ostc_set_field(self.ctx, b, field, _dp(a)))

    def set_geometry(self, b: int, jac: np.ndarray, metric: np.ndarray, elem_stride=None, comp_stride=None,
                     offset=0):
        jac = np.ascontiguousarray(jac, dtype=np.float64)
        metric = np.ascontiguousarray(metric, dtype=np.float64)
        if elem_stride is None:
            elem_stride, comp_stride = 1, jac.size
        self._chk(lib().ostc_set_geometry(self.ctx, b, _dp(jac), _dp(metric), elem_stride, comp_stride, offset))

    def setup_mixing_layer(self):
        """Uniform grid geometry and the mixing-layer initial state of the case."""
        L, c, g = lib(), self.case, self.ngg
        up = np.array(c["streams"]["upper"], dtype=np.float64)
        lo = np.array(c["streams"]["lower"], dtype=np.float64)
        per = c["periods"]
        for b, blk in enumerate(c["blocks"]):
            shp = padded_shape(blk, g)
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

    def initialize(self):
        self._chk(lib().ostc_initialize(self.ctx))

    def zone(self, b: int) -> Zone:
        z = Zone()
        self._chk(lib().ostc_zone_view(self.ctx, b, C.byref(z)))
        return z

    # -- the reference's operators ---------------------------------------
    def rk3(self, n_steps: int = 1, mode: int = MODE_BATCHED, want_dt: bool = True) -> float:
        dt = C.c_double(0.0)
        self._chk(lib().ostc_step(self.ctx, n_steps, mode, C.byref(dt) if want_dt else None))
        return dt.value

    def set_dual_time(self, d: dict):
        """Arms dual time stepping + DPLUR (temporal_scheme 2); q^(n-1) = the current state."""
        self._chk(lib().ostc_set_dual_time(self.ctx, C.byref(dual_time_struct(d))))

    def dual_time_step(self, n_steps: int = 1) -> int:
        """n physical dual time steps; returns the inner iterations of the next step."""
        it = C.c_int()
        self._chk(lib().ostc_dual_time_step(self.ctx, n_steps, C.byref(it)))
        return it.value

    def compute_rhs(self, b: int):
        self._chk(lib().ostc_stage_rhs(self.ctx, b))

    def compute_rhs_batched(self, b: int):
        self._chk(lib().ostc_batched_rhs(self.ctx, b))

    def global_time_step(self) -> float:
        dt = C.c_double()
        self._chk(lib().ostc_compute_dt(self.ctx, C.byref(dt)))
        return dt.value

    def update_cv_and_bv_rk(self, b: int, rk: int, dt: float):
        self._chk(lib().ostc_stage_update(self.ctx, b, rk, dt))

    def copy_qn(self, b: int):
        self._chk(lib().ostc_copy_qn(self.ctx, b))

    def apply_boundary_conditions(self, b: int):
        self._chk(lib().ostc_apply_bc(self.ctx, b))

    def data_communication(self):
        self._chk(lib().ostc_exchange(self.ctx))

    def update_physical_properties(self, b: int):
        self._chk(lib().ostc_update_properties(self.ctx, b))

    def store_last_step(self):
        self._chk(lib().ostc_store_last_step(self.ctx))

    def compute_residual(self) -> np.ndarray:
        r = np.zeros(4)
        self._chk(lib().ostc_residual(self.ctx, _dp(r)))
        return r

    # -- faces shared with other ranks ------------------------------------
    def parallel_faces(self):
        """[{block, rank, flag_send, flag_recv, offset, count, points}] of the parallel faces."""
        L = lib()
        n = C.c_int()
        self._chk(L.ostc_parallel_face_count(self.ctx, C.byref(n)))
        out = []
        for q in range(n.value):
            info = ParallelInfo()
            self._chk(L.ostc_parallel_face_info(self.ctx, q, C.byref(info)))
            out.append({k: getattr(info, k) for k, _ in ParallelInfo._fields_})
        return out

    def _parallel_total(self):
        t = C.c_longlong()
        self._chk(lib().ostc_parallel_buffers(self.ctx, None, None, C.byref(t)))
        return t.value

    def pack_parallel(self) -> np.ndarray:
        """Packs every parallel face and returns the send buffer (host copy)."""
        self._chk(lib().ostc_pack_parallel(self.ctx))
        out = np.zeros(self._parallel_total())
        self._chk(lib().ostc_parallel_copy(self.ctx, 0, _dp(out)))
        return out

    def unpack_parallel(self, recv: np.ndarray):
        """Copies `recv` (the partners' slices, face by face) into the receive
        buffer and unpacks every parallel face."""
        recv = np.ascontiguousarray(recv, dtype=np.float64)
        if recv.size != self._parallel_total():
            raise ValueError(f"receive buffer of {recv.size} doubles, expected {self._parallel_total()}")
        self._chk(lib().ostc_parallel_copy(self.ctx, 1, _dp(recv)))
        self._chk(lib().ostc_unpack_parallel(self.ctx))

    # -- tuning and timing -------------------------------------------------
    def set_tile(self, index: int, waves_per_cu: int = 0, split: int = 0):
        self._chk(lib().ostc_set_tile(self.ctx, index, waves_per_cu, split))

    def set_ghost_props(self, mode: int):
        """GHOST_PROPS_IN_OPS (default): the ghost ops of a batched stage
        recompute the properties of the nodes they write; GHOST_PROPS_PASS: one
        more launch over the ghost layers."""
        self._chk(lib().ostc_set_ghost_props(self.ctx, mode))

    def tile_config(self):
        i, w, s = C.c_int(), C.c_int(), C.c_int()
        self._chk(lib().ostc_tile_config(self.ctx, C.byref(i), C.byref(w), C.byref(s)))
        return i.value, w.value, s.value

    def autotune(self, reps: int = 3):
        bi, bm = C.c_int(), C.c_double()
        self._chk(lib().ostc_autotune(self.ctx, reps, C.byref(bi), C.byref(bm)))
        return bi.value, bm.value

    def last_step_timing(self):
        ms, n = C.c_double(), C.c_longlong()
        self._chk(lib().ostc_last_step_timing(self.ctx, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def fused_kernel_ms(self):
        """(average ms, max ms, launches) of the fused kernel in the last timed step call."""
        avg, mx, n = C.c_double(), C.c_double(), C.c_longlong()
        self._chk(lib().ostc_fused_kernel_ms(self.ctx, C.byref(avg), C.byref(mx), C.byref(n)))
        return avg.value, mx.value, n.value

    def sync(self):
        self._chk(lib().ostc_sync(self.ctx))

    # -- reverse mode ------------------------------------------------------
    def rhs_vjp(self, b: int, dq_bar: np.ndarray) -> np.ndarray:
        """(d dq / d cv)^T dq_bar at block b's current state, on the padded domain."""
        self.set(b, FIELD_DQ_BAR, dq_bar)
        self._chk(lib().ostc_rhs_vjp(self.ctx, b, None, None))
        return self.get(b, FIELD_CV_BAR)

    def perturb_fourier(self, b: int, modes: dict, amp=None):
        self._chk(lib().ostc_perturb_fourier(self.ctx, b, C.byref(fourier_struct(modes, amp))))

    def momentum_thickness(self, b: int, theta: dict) -> float:
        v = C.c_double()
        self._chk(lib().ostc_momentum_thickness(self.ctx, b, C.byref(theta_struct(theta)), C.byref(v)))
        return v.value

    def momentum_thickness_noalloc(self, b: int, theta: dict) -> float:
        """momentum_thickness without device allocation (the library's
        static device scratch; blocks of at most 16384 rows)."""
        v = C.c_double()
        self._chk(lib().ostc_momentum_thickness_noalloc(self.ctx, b, C.byref(theta_struct(theta)), C.byref(v)))
        return v.value

    def adjoint_forward(self, n_steps: int, checkpoint_every: int, theta: dict) -> float:
        v = C.c_double()
        self._chk(lib().ostc_adjoint_forward(self.ctx, n_steps, checkpoint_every, C.byref(theta_struct(theta)),
                                             C.byref(v)))
        return v.value

    def adjoint_backward(self, modes: dict | None = None, amp=None) -> np.ndarray | None:
        if modes is None:
            self._chk(lib().ostc_adjoint_backward(self.ctx, None, None))
            return None
        f = fourier_struct(modes, amp)
        g = np.zeros(f.n_modes)
        self._chk(lib().ostc_adjoint_backward(self.ctx, C.byref(f), _dp(g)))
        return g

    def adjoint_timing(self):
        fwd, bwd = C.c_double(), C.c_double()
        self._chk(lib().ostc_adjoint_timing(self.ctx, C.byref(fwd), C.byref(bwd)))
        return fwd.value, bwd.value

    def setup_fourier_case(self):
        """Geometry, the unperturbed mixing layer and the Fourier perturbation
        of a differentiation case (cases.fourier_case), then initialisation."""
        self.setup_mixing_layer()
        self.perturb_fourier(0, self.case["modes"])
        self.initialize()


def tile_shapes():
    L = lib()
    n = C.c_int()
    L.ostc_tile_count(C.byref(n))
    out = []
    for i in range(n.value):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        L.ostc_tile_shape(i, C.byref(a), C.byref(b), C.byref(c))
        out.append((a.value, b.value, c.value))
    return out
