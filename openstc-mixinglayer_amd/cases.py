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
BC_FARFIELD, BC_OUTFLOW, BC_PERIODIC = 4, 6, 8


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


def mixing_layer_case(x_cells, y_cells, z_cells, *, re_delta=500.0, ma_c=0.4,
                      lx_over_delta=20.0, ly_over_delta=40.0, lz_over_delta=20.0,
                      scheme=51, viscous_order=2, cfl=0.5, prandtl=0.72,
                      seed=42, perturbation=1e-3, y_bc=BC_OUTFLOW):
    """Temporal mixing layer of air on a tensor-product partition.

    ``x_cells`` etc. are lists of per-block cell counts along each axis
    (a block with n cells has n+1 nodes; neighbours share the interface
    node).  x and z are periodic with period = total cells (node N-1 is
    node 0); y uses ``y_bc`` on both ends.  Re_delta = rho*dU*delta/mu
    fixes delta_omega; dU = u_upper - u_lower = 2*Mc*c for equal
    temperatures.
    """
    st = air_streams(ma_c=ma_c, ma_upper=ma_c)
    rho, t = st["upper"][0], st["upper"][5]
    du = st["upper"][1] - st["lower"][1]
    delta = re_delta * sutherland(t) / (rho * du)
    nx, ny, nz = sum(x_cells), sum(y_cells), sum(z_cells)
    dx = lx_over_delta * delta / nx
    dy = ly_over_delta * delta / ny
    dz = lz_over_delta * delta / nz
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
                periodic = axis != 1
                if 0 <= q[axis] < nb[axis]:
                    inner.append(dict(block=b, axis=axis, side=side, target=bidx[tuple(q)]))
                elif periodic and nb[axis] > 1:
                    q[axis] %= nb[axis]
                    inner.append(dict(block=b, axis=axis, side=side, target=bidx[tuple(q)]))
                elif periodic:
                    boundaries.append(dict(block=b, axis=axis, side=side, type=BC_PERIODIC))
                else:
                    boundaries.append(dict(block=b, axis=axis, side=side, type=y_bc))
    # face order inside a block: as registered above (x-, x+, y-, y+, z-, z+)
    return dict(
        n_var=5, inviscid_scheme=scheme, viscous_order=viscous_order, cfl=cfl, prandtl=prandtl,
        fixed_time_step=0, dt=0.0,
        rho_ref=st["upper"][0], v_ref=abs(st["upper"][1]) if abs(st["upper"][1]) >= 1 else st["c_upper"],
        streams=st, delta_omega=delta, dx=dx, dy=dy, dz=dz,
        periods=(nx, 0, nz), seed=seed, amplitude=perturbation * du,
        blocks=blocks, boundaries=boundaries, inner_faces=inner,
        n_cells=sum(b["mx"] * b["my"] * b["mz"] for b in blocks))


def single_block_case(n=64, **kw):
    """Config 'plain single block': n^3 nodes, periodic x/z, outflow y."""
    return mixing_layer_case([n - 1], [n - 1], [n - 1], **kw)


def batched_case(seed=42, **kw):
    """24 uneven blocks (4 x 3 x 2 tensor-product cuts drawn from ``seed``),
    block sizes between 16^3 and 192x128x64 nodes, about 12 M cells."""
    rng = np.random.default_rng(seed)
    xs = [int(v) for v in rng.integers(64, 192, size=4)]
    ys = [int(v) for v in rng.integers(48, 128, size=3)]
    zs = [int(v) for v in rng.integers(32, 64, size=2)]
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
