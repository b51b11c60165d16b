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

CV, BV, QN, DQ, MUL, MACH, JAC, METRIC, DT_LOCAL, X, Y, Z, BV_LAST = range(13)


class _Setup(C.Structure):
    _fields_ = [("n_var", C.c_int), ("inviscid_scheme", C.c_int), ("viscous_order", C.c_int),
                ("fixed_time_step", C.c_int), ("dt", C.c_double), ("cfl", C.c_double),
                ("prandtl", C.c_double), ("rho_ref", C.c_double), ("v_ref", C.c_double),
                ("n_threads", C.c_int)]


class _Streams(C.Structure):
    _fields_ = [("upper", C.c_double * 6), ("lower", C.c_double * 6)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
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
        for name in ("orc_cv_from_bv", "orc_copy_qn", "orc_zero_dq", "orc_inviscid", "orc_viscous",
                     "orc_local_dt", "orc_apply_bc", "orc_properties"):
            getattr(L, name).argtypes = [vp, i]
        for name in ("orc_initialize", "orc_communicate", "orc_store_last_step"):
            getattr(L, name).argtypes = [vp]
        L.orc_global_dt.restype = d
        L.orc_global_dt.argtypes = [vp]
        L.orc_update.argtypes = [vp, i, d, i]
        L.orc_residual.argtypes = [vp, dp]
        for name in ("orc_weno5z", "orc_weno7z", "orc_weno5js"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [dp, dp, d]
        L.orc_lf_split.argtypes = [dp, i, dp, d, dp, dp]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """A set of blocks on the host with the reference's layout."""

    def __init__(self, case: dict, n_threads: int = 0):
        L = lib()
        s = _Setup(case["n_var"], case["inviscid_scheme"], case["viscous_order"],
                   int(case.get("fixed_time_step", 0)), float(case.get("dt", 0.0)), case["cfl"],
                   case["prandtl"], case["rho_ref"], case["v_ref"], n_threads)
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
            L.orc_add_boundary(self.ctx, f["block"], f["axis"], f["side"], f["type"])
        for f in case["inner_faces"]:
            L.orc_add_inner_face(self.ctx, f["block"], f["axis"], f["side"], f["target"])

    def __del__(self):
        try:
            lib().orc_destroy(self.ctx)
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
        """Zero-copy view, shape (n_planes, k, j, i)."""
        n = C.c_longlong()
        p = lib().orc_array(self.ctx, b, which, C.byref(n))
        shp = self.shape(b, which)
        per = shp[0] * shp[1] * shp[2]
        flat = np.ctypeslib.as_array(p, shape=(n.value,))
        return flat.reshape(n.value // per, *shp)

    # -- setup --------------------------------------------------------
    def setup_mixing_layer(self):
        L, c = lib(), self.case
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

    def initialize(self):
        lib().orc_initialize(self.ctx)

    def step(self, n: int = 1) -> float:
        dt = C.c_double()
        lib().orc_step(self.ctx, n, C.byref(dt))
        return dt.value

    # -- stage-level operators ----------------------------------------
    def rhs(self, b: int):
        """dq = 0; inviscid; viscous (src/RK.cuh:113-139)."""
        L = lib()
        L.orc_zero_dq(self.ctx, b)
        L.orc_inviscid(self.ctx, b)
        L.orc_viscous(self.ctx, b)

    def local_dt(self, b: int):
        lib().orc_local_dt(self.ctx, b)

    def global_dt(self) -> float:
        return lib().orc_global_dt(self.ctx)

    def update(self, b: int, dt: float, rk: int):
        lib().orc_update(self.ctx, b, dt, rk)

    def residual(self):
        r = np.zeros(4)
        lib().orc_residual(self.ctx, _dp(r))
        return r


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


def lf_split(q, k3, jac):
    q = np.ascontiguousarray(q, dtype=np.float64)
    k3 = np.ascontiguousarray(k3, dtype=np.float64)
    nv = len(q) - 2
    fp, fm = np.zeros(nv), np.zeros(nv)
    lib().orc_lf_split(_dp(q), nv, _dp(k3), float(jac), _dp(fp), _dp(fm))
    return fp, fm
