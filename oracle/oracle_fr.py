"""ctypes binding of the finite-rate mixture oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
import this module.  The oracle runs on tables read by `mech.load` (this
directory's own reader), never by the product's reader.  Geometry and the
velocity/thermodynamic initial state come from the air oracle's routines
(uniform grid metric, seeded tanh profile), species from the two streams.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

import mech
import oracle as OA

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libostc_oracle_fr.so")
R_U = 8314.462618

(CV, BV, SV, QN, DQ, MUL, LAMBDA, CP, GAMMA, C_SOUND, MACH, RHOD, JAC, METRIC, DT_LOCAL) = range(15)


class _Setup(C.Structure):
    _fields_ = [("inviscid_scheme", C.c_int), ("viscous_order", C.c_int), ("reaction", C.c_int),
                ("positive_preserving", C.c_int), ("grad_p_diffusion", C.c_int), ("fixed_time_step", C.c_int),
                ("reconstruction", C.c_int), ("dt", C.c_double), ("cfl", C.c_double), ("prandtl", C.c_double),
                ("rho_ref", C.c_double), ("v_ref", C.c_double), ("n_threads", C.c_int),
                ("entropy_fix_factor", C.c_double), ("ch_stencil", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            OA.build(force=True)
        L = C.CDLL(LIB_PATH)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        dp, tp = C.POINTER(C.c_double), C.POINTER(mech.MechTables)
        L.orf_create.restype = vp
        L.orf_create.argtypes = [C.POINTER(_Setup), tp]
        L.orf_destroy.argtypes = [vp]
        L.orf_add_block.argtypes = [vp, i, i, i]
        L.orf_ngg.argtypes = [vp]
        L.orf_add_boundary.argtypes = [vp, i, i, i, i]
        L.orf_add_inner_face.argtypes = [vp, i, i, i, i]
        L.orf_add_inner_face_range.argtypes = [vp, i, C.POINTER(C.c_int), C.POINTER(C.c_int), i]
        L.orf_add_boundary_range.argtypes = [vp, i, C.POINTER(C.c_int), i, C.c_void_p]
        L.orf_add_inflow.argtypes = [vp, i, C.POINTER(C.c_int), C.POINTER(OA.Inflow), C.POINTER(C.c_double)]
        L.orf_set_sponge.argtypes = [vp, i, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
        L.orf_array.restype = dp
        L.orf_array.argtypes = [vp, i, i, C.POINTER(C.c_longlong)]
        L.orf_initialize.argtypes = [vp]
        L.orf_step.argtypes = [vp, i, dp]
        for n in ("orf_cv_from_bv", "orf_properties", "orf_zero_dq", "orf_source", "orf_viscous", "orf_local_dt",
                  "orf_apply_bc"):
            getattr(L, n).argtypes = [vp, i]
        L.orf_inviscid.argtypes = [vp, i, d]
        L.orf_global_dt.restype = d
        L.orf_global_dt.argtypes = [vp]
        L.orf_update.argtypes = [vp, i, d, i]
        L.orf_communicate.argtypes = [vp]
        L.orf_enthalpy_cp.argtypes = [tp, d, dp, dp]
        L.orf_gibbs_rt.argtypes = [tp, d, dp]
        L.orf_chemical_source.argtypes = [tp, d, d, dp, dp, dp, dp]
        L.orf_transport.argtypes = [tp, d, d, d, dp, dp, dp, dp]
        L.orf_temperature.restype = d
        L.orf_temperature.argtypes = [tp, d, dp, d, C.POINTER(C.c_int)]
        L.orf_char_matrices.argtypes = [tp, dp, dp, dp, dp, d, d, dp, dp]
        L.orf_riemann.argtypes = [tp, i, dp, dp, dp, dp, d, d, dp]
        L.orf_roe.argtypes = [tp, dp, dp, dp, dp, d, d, d, d, dp]
        L.orf_reconstruct.argtypes = [dp, i, i, dp, dp]
        L.orf_farfield_state.argtypes = [tp, C.POINTER(OA.FarfieldSpec), dp, C.POINTER(OA.Farfield)]
        L.orf_add_farfield.argtypes = [vp, i, i, i, C.POINTER(OA.Farfield)]
        L.orf_store_last_step.argtypes = [vp]
        L.orf_residual.argtypes = [vp, dp]
        _lib = L
    return _lib


class Sponge(C.Structure):
    """orf_sponge (oracle/ostc_oracle_fr.h)."""
    _fields_ = [("enabled", C.c_int * 6), ("start", C.c_double * 6), ("end", C.c_double * 6),
                ("sigma", C.c_double * 6), ("function", C.c_int), ("jacobian_scaled", C.c_int)]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def tables(mech_file, therm_file, tran_file, reaction=1):
    return mech.load(mech_file, therm_file, tran_file, reaction)


# ----------------------------------------------------------- point routines
def farfield_state(T, y, **kw):
    """FarField constructor with species: mass fractions y, keyword inputs
    (mach, pressure, temperature, ...) as for air."""
    yy = np.zeros(mech.MAX_SPEC)
    yy[:T.n_spec] = y
    out = OA.Farfield()
    lib().orf_farfield_state(C.byref(T), C.byref(OA.farfield_spec(**kw)), _dp(yy), C.byref(out))
    return out


def enthalpy_cp(T, t):
    h, cp = np.zeros(mech.MAX_SPEC), np.zeros(mech.MAX_SPEC)
    lib().orf_enthalpy_cp(C.byref(T), t, _dp(h), _dp(cp))
    return h[:T.n_spec], cp[:T.n_spec]


def gibbs_rt(T, t):
    g = np.zeros(mech.MAX_SPEC)
    lib().orf_gibbs_rt(C.byref(T), t, _dp(g))
    return g[:T.n_spec]


def chemical_source(T, rho, t, y):
    y = np.ascontiguousarray(y, dtype=np.float64)
    om, kf, kb = np.zeros(mech.MAX_SPEC), np.zeros(mech.MAX_REAC), np.zeros(mech.MAX_REAC)
    lib().orf_chemical_source(C.byref(T), rho, t, _dp(y), _dp(om), _dp(kf), _dp(kb))
    return om[:T.n_spec], kf[:T.n_reac], kb[:T.n_reac]


def transport(T, t, p, rho, y):
    y = np.ascontiguousarray(y, dtype=np.float64)
    mu, lam, rd = np.zeros(1), np.zeros(1), np.zeros(mech.MAX_SPEC)
    lib().orf_transport(C.byref(T), t, p, rho, _dp(y), _dp(mu), _dp(lam), _dp(rd))
    return mu[0], lam[0], rd[:T.n_spec]


def temperature(T, e, y, t_start):
    y = np.ascontiguousarray(y, dtype=np.float64)
    it = C.c_int(0)
    t = lib().orf_temperature(C.byref(T), e, _dp(y), t_start, C.byref(it))
    return t, it.value


def char_matrices(T, ql, qr, kl, kr, jl, jr):
    nv = 5 + T.n_spec
    L, R = np.zeros(nv * nv), np.zeros(nv * nv)
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (ql, qr, kl, kr)]
    lib().orf_char_matrices(C.byref(T), *[_dp(a) for a in args], jl, jr, _dp(L), _dp(R))
    return L.reshape(nv, nv), R.reshape(nv, nv)


def riemann(T, scheme, ql, qr, kl, kr, jl, jr):
    """AUSM+ (3) or HLLC (4) flux between primitive states [rho,u,v,w,p,Y]."""
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (ql, qr, kl, kr)]
    fc = np.zeros(5 + T.n_spec)
    lib().orf_riemann(C.byref(T), scheme, *[_dp(a) for a in args], jl, jr, _dp(fc))
    return fc


def roe(T, ql, qr, kl, kr, jl, jr, delta_l=0.0, delta_r=0.0):
    """Roe flux with the entropy fix (scheme 2) between primitive states."""
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (ql, qr, kl, kr)]
    fc = np.zeros(5 + T.n_spec)
    lib().orf_roe(C.byref(T), *[_dp(a) for a in args], jl, jr, delta_l, delta_r, _dp(fc))
    return fc


def reconstruct(pv, recon):
    """Left/right states of the face between points 1 and 2 of pv[4, nv]."""
    pv = np.ascontiguousarray(pv, dtype=np.float64)
    nv = pv.shape[1]
    pl, pr = np.zeros(nv), np.zeros(nv)
    lib().orf_reconstruct(_dp(pv), nv, recon, _dp(pl), _dp(pr))
    return pl, pr


def mixture_streams(T, case):
    """Stream states of a two-stream mixture (get_mixing_layer_info with
    species, src/MixingLayer.cu:35-110): [rho,u,v,w,p,T] and Y of the upper
    (fuel) and lower (oxidizer) streams, upper_faster with ma_upper = ma_c."""
    out = {}
    cs = {}
    for side in ("upper", "lower"):
        y = np.array(case["y_" + side], dtype=np.float64)
        t, p = case["t_" + side], case["p_" + side]
        _, cp = enthalpy_cp(T, t)
        mw_inv = 0.0
        cpm = 0.0
        for i in range(T.n_spec):
            mw_inv += y[i] / T.mw[i]
            cpm += cp[i] * y[i]
        gam = cpm / (cpm - R_U * mw_inv)
        cs[side] = math.sqrt(gam * R_U * mw_inv * t)
        out[side] = [p / (R_U * mw_inv * t), 0.0, 0.0, 0.0, p, t]
    ma_c = case["ma_c"]
    ma_l = (ma_c - ma_c) * cs["upper"] / cs["lower"] - ma_c
    out["upper"][1] = ma_c * cs["upper"]
    out["lower"][1] = ma_l * cs["lower"]
    return out


class FrOracle:
    """A set of blocks of the finite-rate mixture path on the host."""

    def __init__(self, case, T=None, n_threads=0):
        self.case = case
        n_threads = n_threads or OA.default_threads()
        self.n_threads = n_threads
        self.T = T if T is not None else tables(*case["mech_files"], reaction=case["reaction"])
        self.L = L = lib()
        s = _Setup(case["inviscid_scheme"], case["viscous_order"], case["reaction"],
                   int(case.get("positive_preserving", 0)), int(case.get("grad_p_diffusion", 0)),
                   int(case.get("fixed_time_step", 0)), int(case.get("reconstruction", 2)),
                   float(case.get("dt", 0.0)), case["cfl"],
                   case["prandtl"], case["rho_ref"], case["v_ref"], n_threads,
                   float(case.get("entropy_fix_factor", 0.125)), int(case.get("ch_stencil", 0)))
        self.ctx = L.orf_create(C.byref(s), C.byref(self.T))
        if not self.ctx:
            raise ValueError("unsupported setup")
        self.ngg = L.orf_ngg(self.ctx)
        self.ns = self.T.n_spec
        for b in case["blocks"]:
            L.orf_add_block(self.ctx, b["mx"], b["my"], b["mz"])
        for f in case["boundaries"]:
            if f["type"] == 5:
                blk = case["blocks"][f["block"]]
                g = L.orf_ngg(self.ctx)
                jj = np.arange(-g, blk["my"] + g)
                y = np.ascontiguousarray(np.broadcast_to((blk["y0"] + jj * case["dy"])[None, :, None],
                                                         (blk["mz"] + 2 * g, blk["my"] + 2 * g, blk["mx"] + 2 * g)))
                if "axes" in blk:
                    raise ValueError("inflow faces on blocks with their own frame: give the y array")
                r = L.orf_add_inflow(self.ctx, f["block"], OA._ints(OA.face_record(blk, f)),
                                     C.byref(OA.inflow_struct(f["inflow"])), _dp(y))
                if r != 0:
                    raise ValueError(f"bad inflow face {f}")
                continue
            ff = farfield_state(self.T, f["farfield_y"], **f["farfield"]) if f["type"] == 4 else None
            if "range" in f:
                r = L.orf_add_boundary_range(self.ctx, f["block"], OA._ints(f["range"]), f["type"],
                                             None if ff is None else C.cast(C.byref(ff), C.c_void_p))
            elif ff is not None:
                r = L.orf_add_farfield(self.ctx, f["block"], f["axis"], f["side"], C.byref(ff))
            else:
                r = L.orf_add_boundary(self.ctx, f["block"], f["axis"], f["side"], f["type"])
            if r != 0:
                raise ValueError(f"bad boundary face {f}")
        if "sponge" in case:
            sp = case["sponge"]
            raw = Sponge((C.c_int * 6)(*[int(v) for v in sp["enabled"]]), (C.c_double * 6)(*sp["start"]),
                         (C.c_double * 6)(*sp["end"]), (C.c_double * 6)(*sp["sigma"]), int(sp.get("function", 0)),
                         int(sp.get("jacobian_scaled", 1)))
            g = L.orf_ngg(self.ctx)
            for b, blk in enumerate(case["blocks"]):
                if "axes" in blk:
                    raise ValueError("sponge layers on blocks with their own frame are not set up here")
                kk, jj, ii = np.meshgrid(np.arange(-g, blk["mz"] + g), np.arange(-g, blk["my"] + g),
                                         np.arange(-g, blk["mx"] + g), indexing="ij")
                xyz = [np.ascontiguousarray(blk[o] + idx * case[d]) for o, idx, d in
                       (("x0", ii, "dx"), ("y0", jj, "dy"), ("z0", kk, "dz"))]
                if L.orf_set_sponge(self.ctx, b, C.cast(C.byref(raw), C.c_void_p), *(_dp(a) for a in xyz)) != 0:
                    raise ValueError("bad sponge layer")
        for f in case["inner_faces"]:
            if "range" in f:
                r = L.orf_add_inner_face_range(self.ctx, f["block"], OA._ints(f["range"]),
                                               OA._ints(f["target_range"]), f["target"])
            else:
                r = L.orf_add_inner_face(self.ctx, f["block"], f["axis"], f["side"], f["target"])
            if r != 0:
                raise ValueError(f"bad inner face {f}")

    def close(self):
        if self.ctx:
            self.L.orf_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def array(self, block, which):
        n = C.c_longlong(0)
        ptr = self.L.orf_array(self.ctx, block, which, C.byref(n))
        flat = np.ctypeslib.as_array(ptr, shape=(n.value,))
        b = self.case["blocks"][block]
        g = self.ngg
        if which in (DQ, DT_LOCAL):
            shape = (b["mz"], b["my"], b["mx"])
        else:
            shape = (b["mz"] + 2 * g, b["my"] + 2 * g, b["mx"] + 2 * g)
        return flat.reshape((-1,) + shape)

    def setup_mixing_layer(self):
        """Geometry and the initial state of every block: the air oracle's
        uniform-grid metric and seeded tanh/velocity-perturbation profile on
        the mixture stream states, species from the stream of each node
        (y >= 0 upper), src/Initialize.cu:259-325."""
        case = self.case
        st = case["streams"]
        # an air oracle with the same ghost depth provides grid and profile
        ao = OA.Oracle(dict(case, n_var=5, inviscid_scheme=71 if self.ngg == 4 else 51), n_threads=self.n_threads)
        streams = OA._Streams((C.c_double * 6)(*st["upper"]), (C.c_double * 6)(*st["lower"]))
        g = self.ngg
        for bi, b in enumerate(case["blocks"]):
            ao.L.orc_uniform_grid(ao.ctx, bi, b["x0"], b["y0"], b["z0"], case["dx"], case["dy"], case["dz"])
            ao.L.orc_compute_metric(ao.ctx, bi)
            px, py, pz = case["periods"]
            ao.L.orc_mixing_layer_ic(ao.ctx, bi, C.byref(streams), case["delta_omega"], b["gi0"], b["gj0"],
                                     b["gk0"], px, py, pz, case["seed"], case["amplitude"])
            self.array(bi, JAC)[...] = ao.array(bi, OA.JAC)
            self.array(bi, METRIC)[...] = ao.array(bi, OA.METRIC)
            self.array(bi, BV)[...] = ao.array(bi, OA.BV)
            yv = b["y0"] + (np.arange(-g, b["my"] + g)) * case["dy"]
            sv = self.array(bi, SV)
            up = yv >= 0
            for l in range(self.ns):
                sv[l][:, up, :] = case["y_upper"][l]
                sv[l][:, ~up, :] = case["y_lower"][l]
        del ao

    def initialize(self):
        self.L.orf_initialize(self.ctx)

    def store_last_step(self):
        self.L.orf_store_last_step(self.ctx)

    def residual(self):
        r = np.zeros(4)
        self.L.orf_residual(self.ctx, _dp(r))
        return r

    def step(self, n=1):
        dt = C.c_double(0)
        self.L.orf_step(self.ctx, n, C.byref(dt))
        return dt.value

    def rhs(self, block, dt=0.0):
        """dq of one block at its current state: source, inviscid, viscous."""
        L = self.L
        L.orf_zero_dq(self.ctx, block)
        L.orf_source(self.ctx, block)
        L.orf_inviscid(self.ctx, block, dt)
        L.orf_viscous(self.ctx, block)
        return self.array(block, DQ)
