"""Mixing-layer case specifications (plain data, no device code).

A case is a dict describing blocks, boundary faces, inner faces, stream
states and scheme keys, in the terms of the reference's setup.txt and
boundary/inner files:

* scheme keys follow ``input/setup.txt`` (``inviscid_scheme`` 51 = WENO-Z5
  component-wise, 71 = WENO-Z7 cp, ``viscous_order`` 2, ``cfl``,
  ``prandtl_number``), see src/Parameter.cpp:187-437;
* boundary labels follow the reference (4 farfield, 6 outflow, 8 periodic),
  src/BoundCond.cuh:1488-1678;
* the stream states follow get_mixing_layer_info for air,
  src/MixingLayer.cu:9-190, with ``upper_stream.mach`` given and
  ``upper_faster`` = 1;
* blocks are vertex-centred and abut with one shared node layer, as the
  reference's multi-block meshes do (src/Mesh.cpp:38-176).

The reference ships no temporal-mixing-layer case (its grid file is
stripped), so the box size, the Reynolds-number-based thickness and the
seeded perturbation are this package's choices, stated here.
"""
from __future__ import annotations

import math

import numpy as np

R_U = 8314.462618          # src/Constants.h:19
GAMMA_AIR = 1.4            # src/Constants.h:23
MW_AIR = 28.9647           # src/Constants.h:28
BC_FARFIELD, BC_INFLOW, BC_OUTFLOW, BC_PERIODIC = 4, 5, 6, 8


def sutherland(t: float) -> float:
    """src/Transport.cu:6-9"""
    return 1.7894e-5 * (t / 288.16) ** 1.5 * (288.16 + 110) / (t + 110)


def air_streams(ma_c=0.4, ma_upper=0.4, t_upper=300.0, t_lower=300.0,
                p_upper=101325.0, p_lower=101325.0, upper_faster=True):
    """Stream states [rho,u,v,w,p,T] for air, src/MixingLayer.cu:64-110."""
    c_u = math.sqrt(GAMMA_AIR * R_U / MW_AIR * t_upper)
    c_l = math.sqrt(GAMMA_AIR * R_U / MW_AIR * t_lower)
    rho_u = p_upper * MW_AIR / (R_U * t_upper)
    rho_l = p_lower * MW_AIR / (R_U * t_lower)
    if upper_faster:
        ma_l = (ma_upper - ma_c) * c_u / c_l - ma_c
    else:
        ma_l = (ma_c + ma_upper) * c_u / c_l + ma_c
    return dict(ma_c=ma_c, ma_upper=ma_upper, t_upper=t_upper, t_lower=t_lower,
                p_upper=p_upper, p_lower=p_lower, upper_faster=bool(upper_faster),
                upper=[rho_u, ma_upper * c_u, 0.0, 0.0, p_upper, t_upper],
                lower=[rho_l, ma_l * c_l, 0.0, 0.0, p_lower, t_lower],
                c_upper=c_u, c_lower=c_l)


def _cuts(lengths):
    """Node offsets of consecutive blocks sharing one node layer."""
    off = [0]
    for n in lengths:
        off.append(off[-1] + n)
    return off


def _partition(x_cells, y_cells, z_cells, dx, dy, dz, delta, ly_over_delta, y_bc, x_open=False):
    """Blocks and faces of a tensor-product partition: x and z periodic
    (inner faces across the period when there are several blocks along the
    axis, a periodic face otherwise), y bounded by ``y_bc`` faces; with
    ``x_open`` the x ends are an inflow face (x-) and an outflow face (x+)."""
    ox, oy, oz = _cuts(x_cells), _cuts(y_cells), _cuts(z_cells)
    blocks, bidx = [], {}
    for kb in range(len(z_cells)):
        for jb in range(len(y_cells)):
            for ib in range(len(x_cells)):
                bidx[(ib, jb, kb)] = len(blocks)
                blocks.append(dict(
                    mx=x_cells[ib] + 1, my=y_cells[jb] + 1, mz=z_cells[kb] + 1,
                    gi0=ox[ib], gj0=oy[jb], gk0=oz[kb],
                    x0=ox[ib] * dx, y0=-0.5 * ly_over_delta * delta + oy[jb] * dy, z0=oz[kb] * dz))
    boundaries, inner = [], []
    nb = (len(x_cells), len(y_cells), len(z_cells))
    for (ib, jb, kb), b in bidx.items():
        pos = (ib, jb, kb)
        for axis in range(3):
            for side in (-1, 1):
                q = list(pos)
                q[axis] += side
                periodic = axis == 2 or (axis == 0 and not x_open)
                if 0 <= q[axis] < nb[axis]:
                    inner.append(dict(block=b, axis=axis, side=side, target=bidx[tuple(q)]))
                elif periodic and nb[axis] > 1:
                    q[axis] %= nb[axis]
                    inner.append(dict(block=b, axis=axis, side=side, target=bidx[tuple(q)]))
                elif periodic:
                    boundaries.append(dict(block=b, axis=axis, side=side, type=BC_PERIODIC))
                elif axis == 0:
                    boundaries.append(dict(block=b, axis=axis, side=side, type=BC_INFLOW if side < 0 else BC_OUTFLOW))
                else:
                    boundaries.append(dict(block=b, axis=axis, side=side, type=y_bc))
    return blocks, boundaries, inner


def stream_farfield(state, c):
    """FarField inputs (mach, pressure, temperature, velocity direction) of a
    stream state [rho, u, v, w, p, T] moving along x with sound speed c."""
    u = state[1]
    return dict(mach=abs(u) / c, pressure=state[4], temperature=state[5], u=-1.0 if u < 0 else 1.0, v=0.0, w=0.0)


def _attach_inflow(boundaries, st, delta, intensity, seed, sv_upper=(), sv_lower=()):
    """The mixing-layer inflow (inflow_type 2) of the two streams on the
    inflow faces, white noise of ``intensity`` (fluctuation_type 1) when
    it is positive."""
    for f in boundaries:
        if f["type"] == BC_INFLOW:
            f["inflow"] = dict(inflow_type=2, fluctuation_type=1 if intensity > 0 else 0, delta_omega=delta,
                               fluctuation_intensity=intensity, seed=seed,
                               upper=[float(v) for v in st["upper"]], lower=[float(v) for v in st["lower"]],
                               sv_upper=[float(v) for v in sv_upper], sv_lower=[float(v) for v in sv_lower])


def _attach_farfield(boundaries, upper, lower):
    """Farfield faces (label 4) on y take the stream they face: y+ the upper
    stream, y- the lower one."""
    for f in boundaries:
        if f["type"] == BC_FARFIELD:
            f["farfield"] = dict(upper if f["side"] > 0 else lower)


def mixing_layer_case(x_cells, y_cells, z_cells, *, re_delta=500.0, ma_c=0.4,
                      lx_over_delta=20.0, ly_over_delta=40.0, lz_over_delta=20.0,
                      scheme=51, viscous_order=2, cfl=0.5, prandtl=0.72,
                      seed=42, perturbation=1e-3, y_bc=BC_OUTFLOW, reconstruction=2,
                      spatial=False, inflow_noise=1e-3):
    """Temporal mixing layer of air on a tensor-product partition.

    ``x_cells`` etc. are lists of per-block cell counts along each axis
    (a block with n cells has n+1 nodes; neighbours share the interface
    node).  x and z are periodic with period = total cells (node N-1 is
    node 0); y uses ``y_bc`` on both ends.  Re_delta = rho*dU*delta/mu
    fixes delta_omega; dU = u_upper - u_lower = 2*Mc*c for equal
    temperatures.  ``reconstruction`` (1 first order, 2 MUSCL, 3 NND2)
    applies to the primitive-variable schemes 1/3/4.  With ``y_bc`` =
    BC_FARFIELD the y faces carry the far-field state of the stream they face.
    ``spatial``: the spatially developing layer -- x- is the mixing-layer
    inflow (src/BoundCond.cuh:420-482) with white noise of relative
    intensity ``inflow_noise``, x+ an outflow face, x not periodic.
    """
    st = air_streams(ma_c=ma_c, ma_upper=ma_c)
    rho, t = st["upper"][0], st["upper"][5]
    du = st["upper"][1] - st["lower"][1]
    delta = re_delta * sutherland(t) / (rho * du)
    nx, ny, nz = sum(x_cells), sum(y_cells), sum(z_cells)
    dx = lx_over_delta * delta / nx
    dy = ly_over_delta * delta / ny
    dz = lz_over_delta * delta / nz
    blocks, boundaries, inner = _partition(x_cells, y_cells, z_cells, dx, dy, dz, delta, ly_over_delta, y_bc,
                                           x_open=spatial)
    _attach_farfield(boundaries, stream_farfield(st["upper"], st["c_upper"]),
                     stream_farfield(st["lower"], st["c_lower"]))
    _attach_inflow(boundaries, st, delta, inflow_noise, seed + 1)
    # face order inside a block: as registered above (x-, x+, y-, y+, z-, z+)
    return dict(
        n_var=5, inviscid_scheme=scheme, viscous_order=viscous_order, cfl=cfl, prandtl=prandtl,
        fixed_time_step=0, dt=0.0, reconstruction=reconstruction,
        rho_ref=st["upper"][0], v_ref=abs(st["upper"][1]) if abs(st["upper"][1]) >= 1 else st["c_upper"],
        streams=st, delta_omega=delta, dx=dx, dy=dy, dz=dz,
        periods=(0 if spatial else nx, 0, nz), seed=seed, amplitude=perturbation * du,
        blocks=blocks, boundaries=boundaries, inner_faces=inner,
        n_cells=sum(b["mx"] * b["my"] * b["mz"] for b in blocks))


def single_block_case(n=64, **kw):
    """Config 'plain single block': n^3 nodes, periodic x/z, outflow y."""
    return mixing_layer_case([n - 1], [n - 1], [n - 1], **kw)


BATCHED_TARGET_CELLS = 12.4e6   # about 480 x 270 x 96 nodes


def _batched_cuts(seed):
    rng = np.random.default_rng(seed)
    xs = [int(v) for v in rng.integers(16, 193, size=4)]
    ys = [int(v) for v in rng.integers(16, 129, size=3)]
    zs = [int(v) for v in rng.integers(16, 65, size=2)]
    return xs, ys, zs


def batched_seed(first=42, rel=0.03):
    """The first seed from ``first`` upward whose 4 x 3 x 2 cuts total
    BATCHED_TARGET_CELLS nodes to within ``rel`` (seed 66: 12,493,728)."""
    seed = first
    while True:
        xs, ys, zs = _batched_cuts(seed)
        n = sum((x + 1) * (y + 1) * (z + 1) for x in xs for y in ys for z in zs)
        if abs(n - BATCHED_TARGET_CELLS) <= rel * BATCHED_TARGET_CELLS:
            return seed
        seed += 1


def batched_case(seed=None, **kw):
    """24 uneven blocks: 4 x 3 x 2 tensor-product cuts of 16..192 (x),
    16..128 (y) and 16..64 (z) cells drawn from ``seed`` (default: the
    first seed from 42 whose total is within 3 % of 12.4 M nodes)."""
    xs, ys, zs = _batched_cuts(batched_seed() if seed is None else seed)
    return mixing_layer_case(xs, ys, zs, **kw)


def small_multiblock_case(seed=7, **kw):
    """A 3 x 2 x 2 partition of a small box for parity tests on the CPU."""
    rng = np.random.default_rng(seed)
    xs = [int(v) for v in rng.integers(6, 14, size=3)]
    ys = [int(v) for v in rng.integers(8, 16, size=2)]
    zs = [int(v) for v in rng.integers(6, 12, size=2)]
    return mixing_layer_case(xs, ys, zs, **kw)


def fourier_modes(case, n_modes=32, seed=3, scale=1e-2, centre=0.25):
    """The seeded Fourier modes of the differentiation workload.

    Mode m = 3 q + c perturbs velocity component c (u, v, w) by
    a_m * du * exp(-((y - yc)/delta_omega)^2) * (sin | cos)(2 pi (kx i dx / lx + kz k dz / lz)),
    kx = q % 4, kz = q // 4, du = scale * (u_upper - u_lower),
    yc = centre * delta_omega; lx, lz are the periodic box lengths, so every
    mode is periodic with the box.  Modes with kx = kz = 0 use cos: they
    change the mean profile, and the off-centre envelope makes that move the
    momentum thickness at first order; the others alternate sin and cos.
    The amplitudes a_m ~ U(-1, 1) are the parameters."""
    blk = case["blocks"][0]
    m = np.arange(n_modes)
    q, comp = m // 3, m % 3
    kx, kz = q % 4, q // 4
    trig = np.where((kx == 0) & (kz == 0), 1, (q + comp) % 2)
    rng = np.random.default_rng(seed)
    st = case["streams"]
    return dict(comp=comp, kx=kx, kz=kz, trig=trig, amp=rng.uniform(-1.0, 1.0, n_modes),
                y0=blk["y0"], dx=case["dx"], dy=case["dy"], dz=case["dz"],
                lx=(blk["mx"] - 1) * case["dx"], lz=(blk["mz"] - 1) * case["dz"],
                delta=case["delta_omega"], yc=centre * case["delta_omega"],
                du=scale * (st["upper"][1] - st["lower"][1]))


def fixed_dt(case, safety=0.9):
    """CFL time step of the unperturbed layer (largest |u|, uniform c and mu),
    times ``safety``: the spectral radii of local_time_step,
    src/TimeAdvanceFunc.cuh:28-83, on a uniform grid."""
    st = case["streams"]
    rho, t = st["upper"][0], st["upper"][5]
    c = math.sqrt(GAMMA_AIR * R_U / MW_AIR * t)
    u = max(abs(st["upper"][1]), abs(st["lower"][1]))
    dx, dy, dz = case["dx"], case["dy"], case["dz"]
    sr_inv = (u + c) / dx + c / dy + c / dz
    sr_visc = (1 / dx ** 2 + 1 / dy ** 2 + 1 / dz ** 2) * max(GAMMA_AIR, 4 / 3) / rho * sutherland(t) / case["prandtl"]
    return safety * case["cfl"] / (sr_inv + sr_visc)


def fourier_case(n=128, n_steps=50, checkpoint_every=10, seed=3, n_modes=32, scheme=51, **kw):
    """Differentiation workload: the n^3 single-block temporal mixing layer
    (air, WENO-Z5 + LF, viscous) with the seeded random perturbation replaced
    by ``n_modes`` Fourier modes whose amplitudes are the parameters, a fixed
    time step, and the momentum thickness of the final state as objective."""
    case = single_block_case(n, scheme=scheme, perturbation=0.0, **kw)
    case["fixed_time_step"] = 1
    case["dt"] = fixed_dt(case)
    case["modes"] = fourier_modes(case, n_modes=n_modes, seed=seed)
    st = case["streams"]
    blk = case["blocks"][0]
    case["theta"] = dict(u1=st["upper"][1], u2=st["lower"][1], rho0=st["upper"][0],
                         du=st["upper"][1] - st["lower"][1], dy=case["dy"], ni=blk["mx"] - 1, nk=blk["mz"] - 1)
    case["n_steps"] = n_steps
    case["checkpoint_every"] = checkpoint_every
    return case


# ------------------------------------------------------ finite-rate mixtures
def nasa_cp(tables, t):
    """Species cp [J/(kg K)] at t (compute_cp, Combustion2Part,
    src/ChemData.cpp:150-176) from mechanism tables."""
    ns = tables.n_spec
    cp = np.zeros(ns)
    for i in range(ns):
        tt = max(t, tables.t_low[i])
        c = tables.low_coeff[i] if tt < tables.t_mid[i] else tables.high_coeff[i]
        cp[i] = (c[0] + c[1] * tt + c[2] * tt ** 2 + c[3] * tt ** 3 + c[4] * tt ** 4) * R_U / tables.mw[i]
    return cp


def mixture_viscosity(tables, t, y):
    """Wilke mixture viscosity of the host routine compute_viscosity,
    src/Transport.cu:11-41."""
    ns = tables.n_spec
    mw = np.array(tables.mw[:ns])
    mw_total = 1.0 / np.sum(np.asarray(y) / mw)
    x = np.asarray(y) * mw_total / mw
    vis = np.zeros(ns)
    for i in range(ns):
        t_dl = t * tables.lj_potent_inv[i]
        vis[i] = tables.vis_coeff[i] * math.sqrt(t) / (1.147 * t_dl ** -0.145 + (t_dl + 0.5) ** -2)
    mu = 0.0
    for i in range(ns):
        s = 0.0
        for j in range(ns):
            if i == j:
                phi = 1.0
            else:
                num = 1 + math.sqrt(vis[i] / vis[j]) * tables.wj_div_wi_to_one4th[i][j]
                phi = num * num * tables.sqrt_wi_div_wj_pl1_mul8[i][j]
            s += phi * x[j]
        mu += vis[i] * x[i] / s
    return mu


def mixture_streams(tables, y_upper, y_lower, t_upper, t_lower, p_upper=101325.0, p_lower=101325.0, ma_c=0.5):
    """Stream states [rho,u,v,w,p,T] of a two-stream mixture,
    get_mixing_layer_info with species (src/MixingLayer.cu:35-110),
    ``upper_stream.mach`` = Mc, ``upper_faster`` = 1, so u_lower = -Mc c_lower
    and the convective Mach number is Mc."""
    out = {}
    for side, y, t, p in (("upper", y_upper, t_upper, p_upper), ("lower", y_lower, t_lower, p_lower)):
        y = np.asarray(y, dtype=np.float64)
        mw_inv = float(np.sum(y / np.array(tables.mw[:tables.n_spec])))
        cp = float(np.sum(nasa_cp(tables, t) * y))
        gam = cp / (cp - R_U * mw_inv)
        c = math.sqrt(gam * R_U * mw_inv * t)
        out["c_" + side] = c
        out[side] = [p / (R_U * mw_inv * t), 0.0, 0.0, 0.0, p, t]
    out["upper"][1] = ma_c * out["c_upper"]
    ma_l = (ma_c - ma_c) * out["c_upper"] / out["c_lower"] - ma_c
    out["lower"][1] = ma_l * out["c_lower"]
    return out


def species_vector(tables, composition: dict):
    names = [tables.spec_name[i].value.decode() for i in range(tables.n_spec)]
    y = np.zeros(len(names))
    for k, v in composition.items():
        y[names.index(k)] = v
    return list(y)


FUEL_H2_N2 = {"H2": 0.05, "N2": 0.95}          # fuel stream, 300 K
OXIDIZER_AIR = {"O2": 0.232, "N2": 0.768}      # oxidizer stream, 1500 K


def reacting_case(x_cells, y_cells, z_cells, tables, *, mech_files=None, reaction=1, re_delta=500.0, ma_c=0.5,
                  fuel=FUEL_H2_N2, oxidizer=OXIDIZER_AIR, t_fuel=300.0, t_oxidizer=1500.0,
                  lx_over_delta=20.0, ly_over_delta=40.0, lz_over_delta=20.0, scheme=52, viscous_order=2,
                  cfl=0.5, prandtl=0.72, positive_preserving=0, seed=42, perturbation=1e-3, y_bc=BC_OUTFLOW,
                  spatial=False, inflow_noise=1e-3):
    """Temporal H2/air mixing layer on a tensor-product partition: the fuel
    (H2 in N2) stream above y = 0, the oxidizer (air) stream below, tanh
    velocity profile with convective Mach ``ma_c``, species and temperature
    steps at y = 0 (src/Initialize.cu:259-325), seeded velocity noise.
    delta_omega follows from Re_delta = rho dU delta / mu of the fuel stream.
    ``spatial``: the spatially developing layer -- x- is the mixing-layer
    inflow of the two streams (src/BoundCond.cuh:420-482) with white noise
    of relative intensity ``inflow_noise``, x+ an outflow face; otherwise
    the temporal layer, periodic in x."""
    y_up = species_vector(tables, fuel)
    y_lo = species_vector(tables, oxidizer)
    st = mixture_streams(tables, y_up, y_lo, t_fuel, t_oxidizer, ma_c=ma_c)
    rho = st["upper"][0]
    du = st["upper"][1] - st["lower"][1]
    delta = re_delta * mixture_viscosity(tables, t_fuel, y_up) / (rho * du)
    nx, ny, nz = sum(x_cells), sum(y_cells), sum(z_cells)
    dx = lx_over_delta * delta / nx
    dy = ly_over_delta * delta / ny
    dz = lz_over_delta * delta / nz
    blocks, boundaries, inner = _partition(x_cells, y_cells, z_cells, dx, dy, dz, delta, ly_over_delta, y_bc,
                                           x_open=spatial)
    _attach_farfield(boundaries, stream_farfield(st["upper"], st["c_upper"]),
                     stream_farfield(st["lower"], st["c_lower"]))
    _attach_inflow(boundaries, st, delta, inflow_noise, seed + 1, y_up, y_lo)
    for f in boundaries:
        if f["type"] == BC_FARFIELD:
            f["farfield_y"] = [float(v) for v in (y_up if f["side"] > 0 else y_lo)]
    ns = tables.n_spec
    u1 = abs(st["upper"][1])
    return dict(
        n_var=5 + ns, n_spec=ns, inviscid_scheme=scheme, viscous_order=viscous_order, reaction=reaction,
        positive_preserving=positive_preserving, cfl=cfl, prandtl=prandtl, fixed_time_step=0, dt=0.0,
        rho_ref=rho, v_ref=u1 if u1 >= 1 else st["c_upper"],
        streams=st, y_upper=y_up, y_lower=y_lo, delta_omega=delta, dx=dx, dy=dy, dz=dz,
        periods=(0 if spatial else nx, 0, nz), seed=seed, amplitude=perturbation * du, mech_files=mech_files,
        blocks=blocks, boundaries=boundaries, inner_faces=inner,
        n_cells=sum(b["mx"] * b["my"] * b["mz"] for b in blocks))


def reacting_single_block(nx, ny, nz, tables, **kw):
    """One block of nx x ny x nz nodes, periodic x/z, outflow y."""
    return reacting_case([nx - 1], [ny - 1], [nz - 1], tables, **kw)


# ------------------------------------------------------------------ rotated pair
ROTATED_AXES = ((0, -1, 0), (-1, 0, 0), (0, 0, -1))  # global (i, j, k) from local (p, q, r), reversed


def rotated_pair(case: dict, cut: int, split_k: int | None = None) -> dict:
    """The single-block case ``case`` (periodic x and z, y faces of one type)
    cut at global node i = ``cut`` into two blocks given by the records of
    the reference's connectivity files.

    Block 0 keeps nodes i <= cut on the global index axes.  Block 1 holds
    nodes i >= cut with local (p, q, r) = (ny-1-j, nx-1-i, nz-1-k): index
    axes permuted and all three reversed (a right-handed frame).  Boundary
    faces are boundary N.txt ranges (block 0's z-min face in two pieces whose
    ranges share the node line i = cut // 2, widened by ngg as the reference
    widens them, so both pieces write the ghosts near that line);
    inner faces are inner N.txt line pairs for the interface at i = cut
    (in two pieces split at 1-based k = ``split_k`` when given) and for the
    periodic closure at i = 0 (node nx-1 is node 0).  ``frames`` records
    each block's map to global node indices: global = origin + axes @ local.
    """
    blk = case["blocks"][0]
    nx, ny, nz = blk["mx"], blk["my"], blk["mz"]
    if len(case["blocks"]) != 1 or not 3 <= cut <= nx - 4:
        raise ValueError("rotated_pair takes a single-block case and an interior cut")
    nq = nx - cut
    by_face = {(f["axis"], f["side"]): f for f in case["boundaries"]}
    y_lo, y_hi = by_face[(1, -1)], by_face[(1, 1)]

    def bnd(b, rng, like):
        f = dict(block=b, range=list(rng), type=like["type"])
        for key in ("farfield", "farfield_y"):
            if key in like:
                f[key] = like[key]
        return f

    c1 = cut // 2
    per = dict(type=BC_PERIODIC)
    boundaries = [
        bnd(0, (0, cut, 0, 0, 0, nz - 1), y_lo), bnd(0, (0, cut, ny - 1, ny - 1, 0, nz - 1), y_hi),
        bnd(0, (0, c1, 0, ny - 1, 0, 0), per), bnd(0, (c1, cut, 0, ny - 1, 0, 0), per),
        bnd(0, (0, cut, 0, ny - 1, nz - 1, nz - 1), per),
        # block 1: p = 0 is global y = ny-1 (the upper face), p = ny-1 the lower one
        bnd(1, (0, 0, 0, nq - 1, 0, nz - 1), y_hi), bnd(1, (ny - 1, ny - 1, 0, nq - 1, 0, nz - 1), y_lo),
        bnd(1, (0, ny - 1, 0, nq - 1, 0, 0), per), bnd(1, (0, ny - 1, 0, nq - 1, nz - 1, nz - 1), per),
    ]
    pieces = [(1, nz)] if split_k is None else [(1, split_k), (split_k, nz)]
    inner = [dict(block=0, range=[1, 1, -1, -ny, 1, nz], target_range=[-ny, -1, 1, 1, nz, 1], target=1)]
    for k0, k1 in pieces:
        inner.append(dict(block=0, range=[cut + 1, cut + 1, -1, -ny, k0, k1],
                          target_range=[-ny, -1, nq, nq, nz + 1 - k0, nz + 1 - k1], target=1))
    inner.append(dict(block=1, range=[-ny, -1, 1, 1, nz, 1], target_range=[1, 1, -1, -ny, 1, nz], target=0))
    for k0, k1 in pieces:
        inner.append(dict(block=1, range=[-ny, -1, nq, nq, nz + 1 - k0, nz + 1 - k1],
                          target_range=[cut + 1, cut + 1, -1, -ny, k0, k1], target=0))
    blocks = [dict(blk, mx=cut + 1, my=ny, mz=nz, origin=(0, 0, 0), axes=((1, 0, 0), (0, 1, 0), (0, 0, 1))),
              dict(blk, mx=ny, my=nq, mz=nz, origin=(nx - 1, ny - 1, nz - 1), axes=ROTATED_AXES)]
    out = dict(case)
    out.update(blocks=blocks, boundaries=boundaries, inner_faces=inner,
               n_cells=sum(b["mx"] * b["my"] * b["mz"] for b in blocks), global_shape=(nx, ny, nz))
    return out


def _global_index(blk: dict, p, q, r):
    o = blk.get("origin", (0, 0, 0))
    a = blk.get("axes", ((1, 0, 0), (0, 1, 0), (0, 0, 1)))
    return tuple(o[d] + a[d][0] * p + a[d][1] * q + a[d][2] * r for d in range(3))


def _local_grid(blk: dict, g: int):
    p = np.arange(-g, blk["mx"] + g)
    q = np.arange(-g, blk["my"] + g)
    r = np.arange(-g, blk["mz"] + g)
    R, Q, P = np.meshgrid(r, q, p, indexing="ij")
    return _global_index(blk, P, Q, R)


def node_coordinates(case: dict, b: int, g: int):
    """Node coordinates (x, y, z) of block b on its domain padded by g
    layers, arrays (k, j, i), from the uniform grid of the case (a block
    with ``origin``/``axes`` maps its local indices to the grid's first)."""
    blk = case["blocks"][b]
    i, j, k = _local_grid(blk, g)
    return blk["x0"] + i * case["dx"], blk["y0"] + j * case["dy"], blk["z0"] + k * case["dz"]


def block_coordinates(case: dict, b: int, g: int):
    """node_coordinates on the domain padded by g+1 layers (the layers the
    metric's central differences read)."""
    return node_coordinates(case, b, g + 1)


def face_range(blk: dict, f: dict):
    """The boundary N.txt record of a boundary face: its ``range``, or the
    whole face of ``axis``/``side``."""
    if "range" in f:
        return list(f["range"])
    ext = (blk["mx"], blk["my"], blk["mz"])
    r = []
    for d in range(3):
        r += [0, ext[d] - 1]
    a = f["axis"]
    r[2 * a] = r[2 * a + 1] = 0 if f["side"] < 0 else ext[a] - 1
    return r


def block_from_global(case: dict, b: int, arr: np.ndarray, g: int) -> np.ndarray:
    """Block b's padded array (planes, k, j, i) from a global single-block
    padded array with g ghost layers."""
    i, j, k = _local_grid(case["blocks"][b], g)
    return arr[:, k + g, j + g, i + g]


def global_from_block(case: dict, b: int, arr: np.ndarray, g: int):
    """(global (k, j, i) index arrays, values (planes, n)) of block b's
    interior nodes; arr is padded by g (g = 0 for interior-only arrays)."""
    blk = case["blocks"][b]
    i, j, k = _local_grid(blk, 0)
    inner = arr[:, g:g + blk["mz"], g:g + blk["my"], g:g + blk["mx"]] if g else arr
    return (k.ravel(), j.ravel(), i.ravel()), inner.reshape(arr.shape[0], -1)


def sponge_layer(case: dict, sides=(1, 2, 3), depth=0.2, coefficient=1.0, v_char=None,
                 jacobian_scaled=1) -> dict:
    """A sponge layer for ``case`` (setup.txt keys sponge_layer_direction,
    sponge*Start/End, spongeCoefficient): on each listed side (0 x-, 1 x+,
    2 y-, 3 y+, 4 z-, 5 z+) it ramps up over the outer ``depth`` fraction of
    the domain; sigma_d = coefficient * v_char / |start - end|
    (src/DParameter.cu:258-282), v_char defaulting to the velocity
    difference of the two streams.  jacobian_scaled 1 multiplies the term
    by the Jacobian (relaxation rate sigma); 0 keeps the reference's term,
    whose rate is sigma / J."""
    b = case["blocks"]
    lo = [min(blk[o] for blk in b) for o in ("x0", "y0", "z0")]
    hi = [max(blk[o] + (blk[m] - 1) * case[d] for blk in b) for o, m, d in
          (("x0", "mx", "dx"), ("y0", "my", "dy"), ("z0", "mz", "dz"))]
    st = case["streams"]
    v = abs(st["upper"][1] - st["lower"][1]) if v_char is None else v_char
    enabled, start, end, sigma = [0] * 6, [0.0] * 6, [0.0] * 6, [0.0] * 6
    for side in sides:
        a = side // 2
        width = depth * (hi[a] - lo[a])
        if side % 2 == 0:   # minus side: start inside, end at the boundary (start > end)
            start[side], end[side] = lo[a] + width, lo[a]
        else:
            start[side], end[side] = hi[a] - width, hi[a]
        enabled[side] = 1
        sigma[side] = coefficient * v / abs(start[side] - end[side])
    return dict(enabled=enabled, start=start, end=end, sigma=sigma, function=0, jacobian_scaled=jacobian_scaled)
