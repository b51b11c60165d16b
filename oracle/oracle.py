"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
import this module; it is the checker, never the thing measured or shipped.
See ostc_oracle.cpp for what each entry point restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libostc_oracle.so")
TANGENT_LIB_PATH = os.path.join(HERE, "_build", "libostc_oracle_tangent.so")
COUNT_LIB_PATH = os.path.join(HERE, "_build", "libostc_oracle_count.so")

CV, BV, QN, DQ, MUL, MACH, JAC, METRIC, DT_LOCAL, X, Y, Z, BV_LAST = range(13)


class _Setup(C.Structure):
    _fields_ = [("n_var", C.c_int), ("inviscid_scheme", C.c_int), ("viscous_order", C.c_int),
                ("fixed_time_step", C.c_int), ("dt", C.c_double), ("cfl", C.c_double),
                ("prandtl", C.c_double), ("rho_ref", C.c_double), ("v_ref", C.c_double),
                ("n_threads", C.c_int), ("reconstruction", C.c_int), ("entropy_fix_factor", C.c_double)]


class _Streams(C.Structure):
    _fields_ = [("upper", C.c_double * 6), ("lower", C.c_double * 6)]


class Farfield(C.Structure):
    """FarField members after its constructor (src/BoundCond.h:129-149)."""
    _fields_ = [("density", C.c_double), ("u", C.c_double), ("v", C.c_double), ("w", C.c_double),
                ("pressure", C.c_double), ("temperature", C.c_double), ("acoustic_speed", C.c_double),
                ("entropy", C.c_double), ("specific_heat_ratio", C.c_double), ("mw", C.c_double),
                ("sv", C.c_double * 16)]


class FarfieldSpec(C.Structure):
    """FarField constructor inputs; -1 = not given (velocity 0, direction (1, 0, 0))."""
    _fields_ = [("mach", C.c_double), ("pressure", C.c_double), ("temperature", C.c_double),
                ("velocity", C.c_double), ("density", C.c_double), ("u", C.c_double), ("v", C.c_double),
                ("w", C.c_double), ("reynolds_number", C.c_double)]


def farfield_spec(**kw) -> FarfieldSpec:
    d = dict(mach=-1.0, pressure=-1.0, temperature=-1.0, velocity=0.0, density=-1.0, u=1.0, v=0.0, w=0.0,
             reynolds_number=-1.0)
    d.update(kw)
    return FarfieldSpec(**{k: float(v) for k, v in d.items()})


def farfield_state_air(**kw) -> Farfield:
    out = Farfield()
    if lib().orc_farfield_state_air(C.byref(farfield_spec(**kw)), C.byref(out)) != 0:
        raise ValueError("bad far-field specification")
    return out


class DualTime(C.Structure):
    """orc_dual_time: the dual time stepping keys of setup.txt."""
    _fields_ = [("dt", C.c_double), ("inner_iteration", C.c_int), ("max_inner_iteration", C.c_int),
                ("iteration_adjust_step", C.c_int), ("dplur_inner_step", C.c_int), ("step", C.c_int)]


def dual_time_struct(d: dict) -> DualTime:
    return DualTime(d["dt"], d.get("inner_iteration", 20), d.get("max_inner_iteration", 30),
                    d.get("iteration_adjust_step", 20), d.get("dplur_inner_step", 3), d.get("step", 0))


class _Fourier(C.Structure):
    _fields_ = [("n_modes", C.c_int), ("comp", C.c_void_p), ("kx", C.c_void_p), ("kz", C.c_void_p),
                ("trig", C.c_void_p), ("amp", C.c_void_p), ("amp_dot", C.c_void_p),
                ("x0", C.c_double), ("z0", C.c_double), ("lx", C.c_double), ("lz", C.c_double),
                ("delta", C.c_double), ("yc", C.c_double), ("du", C.c_double)]


def build(force: bool = False) -> str:
    if force or not all(os.path.exists(p) for p in (LIB_PATH, TANGENT_LIB_PATH, COUNT_LIB_PATH)):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_libs = {}


def default_threads() -> int:
    """OpenMP threads of an oracle created without a count: OMP_NUM_THREADS
    when set, else every core.  The count is set on every creation because
    omp_set_num_threads is process-wide: an oracle made with one thread
    would otherwise leave every later one single-threaded."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def lib(tangent=False):
    """The double oracle, or (tangent=True) the same source built with
    {value, tangent} dual numbers as state values, or (tangent="count") the
    operation-counting build (single-threaded; orc_op_counts)."""
    if tangent not in _libs:
        path = COUNT_LIB_PATH if tangent == "count" else (TANGENT_LIB_PATH if tangent else LIB_PATH)
        if not os.path.exists(path):
            build(force=True)
        L = C.CDLL(path)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        dp = C.POINTER(C.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(_Setup)]
        L.orc_destroy.argtypes = [vp]
        for name in ("orc_add_block",):
            getattr(L, name).argtypes = [vp, i, i, i]
        L.orc_ngg.argtypes = [vp]
        L.orc_add_boundary.argtypes = [vp, i, i, i, i]
        L.orc_add_inner_face.argtypes = [vp, i, i, i, i]
        L.orc_add_boundary_range.argtypes = [vp, i, C.POINTER(C.c_int), i, C.c_void_p]
        L.orc_add_inner_face_range.argtypes = [vp, i, C.POINTER(C.c_int), C.POINTER(C.c_int), i]
        L.orc_add_inflow.argtypes = [vp, i, C.POINTER(C.c_int), C.POINTER(Inflow)]
        L.orc_set_sponge.argtypes = [vp, i, C.c_void_p]
        L.orc_source.argtypes = [vp, i]
        L.orc_add_farfield.argtypes = [vp, i, i, i, C.POINTER(Farfield)]
        L.orc_farfield_state_air.argtypes = [C.POINTER(FarfieldSpec), C.POINTER(Farfield)]
        L.orc_array.restype = dp
        L.orc_array.argtypes = [vp, i, i, C.POINTER(C.c_longlong)]
        L.orc_uniform_grid.argtypes = [vp, i, d, d, d, d, d, d]
        L.orc_compute_metric.argtypes = [vp, i]
        L.orc_mixing_layer_streams.argtypes = [d, d, d, d, d, d, i, C.POINTER(_Streams)]
        L.orc_sutherland.restype = d
        L.orc_sutherland.argtypes = [d]
        L.orc_mixing_layer_ic.argtypes = [vp, i, C.POINTER(_Streams), d, i, i, i, i, i, i,
                                          C.c_uint64, d]
        L.orc_step.argtypes = [vp, i, dp]
        L.orc_dual_time_init.argtypes = [vp, C.POINTER(DualTime)]
        L.orc_dual_time_step.argtypes = [vp, i, C.POINTER(C.c_int)]
        L.orc_jacobian_times_dq.argtypes = [dp, dp, d, dp, dp]
        for name in ("orc_cv_from_bv", "orc_copy_qn", "orc_zero_dq", "orc_inviscid", "orc_viscous",
                     "orc_local_dt", "orc_apply_bc", "orc_properties"):
            getattr(L, name).argtypes = [vp, i]
        for name in ("orc_initialize", "orc_communicate", "orc_store_last_step"):
            getattr(L, name).argtypes = [vp]
        L.orc_global_dt.restype = d
        L.orc_global_dt.argtypes = [vp]
        L.orc_update.argtypes = [vp, i, d, i]
        L.orc_residual.argtypes = [vp, dp]
        if tangent == "count":
            L.orc_op_counts.argtypes = [C.POINTER(C.c_ulonglong)]
        for name in ("orc_weno5z", "orc_weno7z", "orc_weno5js"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [dp, dp, d]
        L.orc_lf_split.argtypes = [dp, i, dp, d, dp, dp]
        L.orc_riemann.argtypes = [i, dp, dp, dp, dp, d, d, dp]
        L.orc_roe.argtypes = [dp, dp, dp, dp, d, d, d, d, dp]
        L.orc_char_matrices.argtypes = [dp, dp, dp, dp, d, d, dp, dp]
        L.orc_fourier_perturbation.argtypes = [vp, i, C.POINTER(_Fourier)]
        L.orc_momentum_thickness.restype = d
        L.orc_momentum_thickness.argtypes = [vp, i, d, d, d, d, d, i, i, dp]
        L.orc_scalar_width.restype = i
        _libs[tangent] = L
    return _libs[tangent]


class Sponge(C.Structure):
    """orc_sponge (oracle/ostc_oracle.h)."""
    _fields_ = [("enabled", C.c_int * 6), ("start", C.c_double * 6), ("end", C.c_double * 6),
                ("sigma", C.c_double * 6), ("function", C.c_int), ("jacobian_scaled", C.c_int)]


class Inflow(C.Structure):
    """orc_inflow (oracle/ostc_oracle.h)."""
    _fields_ = [("inflow_type", C.c_int), ("fluctuation_type", C.c_int), ("delta_omega", C.c_double),
                ("fluctuation_intensity", C.c_double), ("seed", C.c_uint64), ("upper", C.c_double * 6),
                ("lower", C.c_double * 6), ("sv_upper", C.c_double * 16), ("sv_lower", C.c_double * 16)]


def inflow_struct(d: dict) -> Inflow:
    pad = lambda v: list(v) + [0.0] * (16 - len(v))
    return Inflow(int(d.get("inflow_type", 2)), int(d.get("fluctuation_type", 0)), float(d.get("delta_omega", 0.0)),
                  float(d.get("fluctuation_intensity", 0.0)), int(d.get("seed", 0)),
                  (C.c_double * 6)(*d["upper"]), (C.c_double * 6)(*d["lower"]),
                  (C.c_double * 16)(*pad(d.get("sv_upper", []))), (C.c_double * 16)(*pad(d.get("sv_lower", []))))


def face_record(blk: dict, f: dict):
    """The boundary N.txt record of a face: its ``range`` or the whole face."""
    if "range" in f:
        return list(f["range"])
    ext = (blk["mx"], blk["my"], blk["mz"])
    r = []
    for d in range(3):
        r += [0, ext[d] - 1]
    r[2 * f["axis"]] = r[2 * f["axis"] + 1] = 0 if f["side"] < 0 else ext[f["axis"]] - 1
    return r


def _ints(v):
    """A C int[6] of a connectivity record."""
    return (C.c_int * len(v))(*[int(x) for x in v])


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """A set of blocks on the host with the reference's layout."""

    def __init__(self, case: dict, n_threads: int = 0, tangent: bool = False):
        self.tangent = tangent
        self.L = L = lib(tangent)
        n_threads = n_threads or default_threads()
        self.n_threads = n_threads
        s = _Setup(case["n_var"], case["inviscid_scheme"], case["viscous_order"],
                   int(case.get("fixed_time_step", 0)), float(case.get("dt", 0.0)), case["cfl"],
                   case["prandtl"], case["rho_ref"], case["v_ref"], n_threads, int(case.get("reconstruction", 2)),
                   float(case.get("entropy_fix_factor", 0.125)))
        self.case = case
        self.ctx = L.orc_create(C.byref(s))
        self.ngg = L.orc_ngg(self.ctx)
        self.nblocks = 0
        for b in case["blocks"]:
            bid = L.orc_add_block(self.ctx, b["mx"], b["my"], b["mz"])
            if bid < 0:
                raise ValueError("bad block")
            self.nblocks += 1
        for f in case["boundaries"]:
            if f["type"] == 5:
                r = L.orc_add_inflow(self.ctx, f["block"], _ints(face_record(case["blocks"][f["block"]], f)),
                                     C.byref(inflow_struct(f["inflow"])))
                if r != 0:
                    raise ValueError(f"bad inflow face {f}")
                continue
            ff = farfield_state_air(**f["farfield"]) if f["type"] == 4 else None
            if "range" in f:
                r = L.orc_add_boundary_range(self.ctx, f["block"], _ints(f["range"]), f["type"],
                                             None if ff is None else C.cast(C.byref(ff), C.c_void_p))
            elif ff is not None:
                r = L.orc_add_farfield(self.ctx, f["block"], f["axis"], f["side"], C.byref(ff))
            else:
                r = L.orc_add_boundary(self.ctx, f["block"], f["axis"], f["side"], f["type"])
            if r != 0:
                raise ValueError(f"bad boundary face {f}")
        if "sponge" in case:
            sp = case["sponge"]
            st = Sponge((C.c_int * 6)(*[int(v) for v in sp["enabled"]]), (C.c_double * 6)(*sp["start"]),
                        (C.c_double * 6)(*sp["end"]), (C.c_double * 6)(*sp["sigma"]), int(sp.get("function", 0)),
                        int(sp.get("jacobian_scaled", 1)))
            for b in range(len(case["blocks"])):
                L.orc_set_sponge(self.ctx, b, C.cast(C.byref(st), C.c_void_p))
        for f in case["inner_faces"]:
            if "range" in f:
                r = L.orc_add_inner_face_range(self.ctx, f["block"], _ints(f["range"]), _ints(f["target_range"]),
                                               f["target"])
            else:
                r = L.orc_add_inner_face(self.ctx, f["block"], f["axis"], f["side"], f["target"])
            if r != 0:
                raise ValueError(f"bad inner face {f}")

    def __del__(self):
        try:
            self.L.orc_destroy(self.ctx)
        except Exception:
            pass

    # -- arrays -------------------------------------------------------
    def shape(self, b: int, which: int):
        blk = self.case["blocks"][b]
        g = self.ngg
        if which in (DQ, DT_LOCAL, BV_LAST):
            g = 0
        if which in (X, Y, Z):
            g = self.ngg + 1
        return (blk["mz"] + 2 * g, blk["my"] + 2 * g, blk["mx"] + 2 * g)

    def array(self, b: int, which: int) -> np.ndarray:
        """Zero-copy view, shape (n_planes, k, j, i); in the tangent build the
        state arrays carry a last axis of 2 (value, tangent)."""
        n = C.c_longlong()
        p = self.L.orc_array(self.ctx, b, which, C.byref(n))
        shp = self.shape(b, which)
        per = shp[0] * shp[1] * shp[2]
        w = 2 if self.tangent is True and which not in (JAC, METRIC, X, Y, Z) else 1
        flat = np.ctypeslib.as_array(p, shape=(n.value,))
        if w == 1:
            return flat.reshape(n.value // per, *shp)
        return flat.reshape(n.value // (per * 2), *shp, 2)

    # -- setup --------------------------------------------------------
    def setup_mixing_layer(self):
        L, c = self.L, self.case
        st = _Streams()
        for i in range(6):
            st.upper[i] = c["streams"]["upper"][i]
            st.lower[i] = c["streams"]["lower"][i]
        self._streams = st
        per = c["periods"]
        for b, blk in enumerate(c["blocks"]):
            L.orc_uniform_grid(self.ctx, b, blk["x0"], blk["y0"], blk["z0"], c["dx"], c["dy"], c["dz"])
            if L.orc_compute_metric(self.ctx, b) != 0:
                raise RuntimeError("negative Jacobian")
            L.orc_mixing_layer_ic(self.ctx, b, C.byref(st), c["delta_omega"], blk["gi0"], blk["gj0"],
                                  blk["gk0"], per[0], per[1], per[2], c["seed"], c["amplitude"])

    def perturb_fourier(self, b: int, modes: dict, amp=None, amp_dot=None):
        """Fourier-mode velocity perturbation (cases.fourier_modes); in the
        tangent build amp_dot seeds the amplitude tangents."""
        a = np.ascontiguousarray(modes["amp"] if amp is None else amp, dtype=np.float64)
        n = len(a)
        ints = [np.ascontiguousarray(modes[k][:n], dtype=np.int32) for k in ("comp", "kx", "kz", "trig")]
        ad = None if amp_dot is None else np.ascontiguousarray(amp_dot, dtype=np.float64)
        blk = self.case["blocks"][b]
        f = _Fourier(n, *(x.ctypes.data for x in ints), a.ctypes.data, None if ad is None else ad.ctypes.data,
                     blk["x0"], blk["z0"], modes["lx"], modes["lz"], modes["delta"], modes["yc"], modes["du"])
        if self.L.orc_fourier_perturbation(self.ctx, b, C.byref(f)) != 0:
            raise ValueError("bad Fourier modes")

    def momentum_thickness(self, b: int, t: dict):
        """(theta, d theta) -- the tangent is 0 in the double build."""
        dth = C.c_double()
        th = self.L.orc_momentum_thickness(self.ctx, b, t["u1"], t["u2"], t["rho0"], t["du"], t["dy"], t["ni"],
                                           t["nk"], C.byref(dth))
        return th, dth.value

    def initialize(self):
        self.L.orc_initialize(self.ctx)

    def step(self, n: int = 1) -> float:
        dt = C.c_double()
        self.L.orc_step(self.ctx, n, C.byref(dt))
        return dt.value

    def dual_time_init(self, d: dict):
        """Dual time stepping + DPLUR settings; q^(n-1) = the current cv."""
        if self.L.orc_dual_time_init(self.ctx, C.byref(dual_time_struct(d))) != 0:
            raise ValueError("bad dual time settings")

    def dual_time_step(self, n: int = 1) -> int:
        """n physical steps; returns the inner iterations of the next step."""
        it = C.c_int()
        r = self.L.orc_dual_time_step(self.ctx, n, C.byref(it))
        if r < 0:
            raise FloatingPointError("NaN in the inner-iteration residual" if r == -1 else "not initialised")
        return it.value

    # -- stage-level operators ----------------------------------------
    def rhs(self, b: int):
        """dq = 0; source (the sponge layer); inviscid; viscous (src/RK.cuh:113-139)."""
        L = self.L
        L.orc_zero_dq(self.ctx, b)
        L.orc_source(self.ctx, b)
        L.orc_inviscid(self.ctx, b)
        L.orc_viscous(self.ctx, b)

    def local_dt(self, b: int):
        self.L.orc_local_dt(self.ctx, b)

    def global_dt(self) -> float:
        return self.L.orc_global_dt(self.ctx)

    def update(self, b: int, dt: float, rk: int):
        self.L.orc_update(self.ctx, b, dt, rk)

    def residual(self):
        r = np.zeros(4)
        self.L.orc_residual(self.ctx, _dp(r))
        return r


def jacobian_times_dq(k3, pv, lam, dq):
    """compute_jacobian_times_dq<Air> at one point."""
    k3 = np.ascontiguousarray(k3, dtype=np.float64)
    pv = np.ascontiguousarray(pv, dtype=np.float64)
    dq = np.ascontiguousarray(dq, dtype=np.float64)
    out = np.zeros(5)
    lib().orc_jacobian_times_dq(_dp(k3), _dp(pv), lam, _dp(dq), _dp(out))
    return out


def weno5z(vp, vm, eps):
    vp = np.ascontiguousarray(vp, dtype=np.float64)
    vm = np.ascontiguousarray(vm, dtype=np.float64)
    return lib().orc_weno5z(_dp(vp), _dp(vm), float(eps))


def weno7z(vp, vm, eps):
    vp = np.ascontiguousarray(vp, dtype=np.float64)
    vm = np.ascontiguousarray(vm, dtype=np.float64)
    return lib().orc_weno7z(_dp(vp), _dp(vm), float(eps))


def weno5js(vp, vm, eps):
    vp = np.ascontiguousarray(vp, dtype=np.float64)
    vm = np.ascontiguousarray(vm, dtype=np.float64)
    return lib().orc_weno5js(_dp(vp), _dp(vm), float(eps))


def riemann(scheme, ql, qr, kl, kr, jl, jr):
    """LF (1) / AUSM+ (3) / HLLC (4) flux of air between primitive states [rho,u,v,w,p]."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (ql, qr, kl, kr)]
    fc = np.zeros(5)
    lib().orc_riemann(scheme, *[_dp(x) for x in a], float(jl), float(jr), _dp(fc))
    return fc


def roe(ql, qr, kl, kr, jl, jr, dl=0.0, dr=0.0):
    """Roe flux of air between primitive states [rho,u,v,w,p] with the two
    points' entropy-fix deltas."""
    def full(q):
        q = np.asarray(q, dtype=np.float64)
        e = 0.5 * q[0] * (q[1] ** 2 + q[2] ** 2 + q[3] ** 2) + q[4] / 0.4
        return np.ascontiguousarray(np.concatenate([q, [e]]))
    a = [full(ql), full(qr)] + [np.ascontiguousarray(x, dtype=np.float64) for x in (kl, kr)]
    fc = np.zeros(5)
    lib().orc_roe(*[_dp(x) for x in a], float(jl), float(jr), float(dl), float(dr), _dp(fc))
    return fc


def char_matrices(ql, qr, kl, kr, jl, jr):
    """Left and right eigenvector matrices of the air characteristic WENO at a
    face between states [cv(5), p]."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (ql, qr, kl, kr)]
    L, R = np.zeros(25), np.zeros(25)
    lib().orc_char_matrices(*[_dp(x) for x in a], float(jl), float(jr), _dp(L), _dp(R))
    return L.reshape(5, 5), R.reshape(5, 5)


def lf_split(q, k3, jac):
    q = np.ascontiguousarray(q, dtype=np.float64)
    k3 = np.ascontiguousarray(k3, dtype=np.float64)
    nv = len(q) - 2
    fp, fm = np.zeros(nv), np.zeros(nv)
    lib().orc_lf_split(_dp(q), nv, _dp(k3), float(jac), _dp(fp), _dp(fm))
    return fp, fm


OP_NAMES = ("add", "mul", "div", "sqrt", "trans")


def op_counts(o) -> dict:
    """Operation counts of a count-build Oracle since the last reset."""
    out = (C.c_ulonglong * 5)()
    o.L.orc_op_counts(out)
    return dict(zip(OP_NAMES, (int(v) for v in out)))


def op_counts_reset(o):
    o.L.orc_op_counts_reset()
