"""Benchmark: FP64 RHS + SSP-RK3 Mcell-updates/s on one MI355X.

Workload: the batched multi-block air mixing layer -- 24 blocks of uneven
sizes (4 x 3 x 2 tensor-product cuts of 16..192 x 16..128 x 16..64 cells),
about 12.5 M cells,
WENO-Z5 component-wise + Lax-Friedrichs splitting, second-order viscous
fluxes, SSP-RK3 at CFL 0.5, in-GPU inner-face exchange.  The BATCHED path
(fused RHS+update kernel over the block work-list, graph-captured steps) is
the measured value; the PLAIN path (per-block launches shaped like the
reference) is timed on the same state beside it.

One update = one full 3-stage SSP-RK3 step of one interior cell.  Inputs are
resident in HBM before the timed region; the timed region is exactly
--steps steps bracketed by device synchronisation.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
pkg = importlib.import_module("openstc-mixinglayer_amd")

METRIC = "FP64 RHS+RK3 Mcell-updates/s vs plain launch; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# Contract-minimal HBM bytes per interior cell per RK stage for air:
# read cv[5] + qn[5] + metric[9] + jac, write cv[5] + bv[6] + mu + Mach.
BYTES_PER_CELL_STAGE = 8 * (5 + 5 + 10) + 8 * (5 + 6 + 2)
# ... and for H2/air (9 species): read cv[14] + qn[14] + metric[9] + jac,
# write cv[14] + bv[6] + sv[9] + mu, lambda, cp, gamma, c, Mach + rhoD[9];
# per step 3 stages plus the qn copy (read and write cv[14]).
BYTES_PER_CELL_STAGE_FR = 8 * (14 + 14 + 10) + 8 * (14 + 6 + 9 + 6 + 9)
BYTES_PER_CELL_STEP_FR = 3 * BYTES_PER_CELL_STAGE_FR + 8 * 14


def cpu_threads():
    """OpenMP threads of the CPU baseline: OMP_NUM_THREADS when the host sets
    it (the share of host cores that goes with one GPU), else every core."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def cpu_model():
    """The host CPU's model name and logical core count (lscpu)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cores": os.cpu_count()}


def cpu_baseline(case, steps: int):
    """The CPU oracle (host C++, OpenMP over cells) on a bounded sample of the
    same workload: `steps` RK3 steps of the full 24-block case."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads = cpu_threads()
    o = O.Oracle(case, n_threads=threads)
    o.setup_mixing_layer()
    o.initialize()
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    cells = case["n_cells"]
    return {"value": cells * steps / dt / 1e6, "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"{steps} RK3 steps of the full {len(case['blocks'])}-block case ({cells} cells), "
                      f"{dt:.1f} s on {threads} OpenMP threads"}


def adjoint_leg(n: int, steps: int, every: int, h=1e-6):
    """Differentiation workload: n^3 single-block mixing layer, 32 Fourier-mode
    amplitudes -> momentum thickness after `steps` fixed-dt SSP-RK3 steps.
    Forward cost = the same steps through the graph-captured batched path;
    adjoint cost = the reverse sweep, recomputation from the checkpoints
    included.  The gradient is checked against central differences of the
    forward run (step h) for every amplitude, norm-wise:
    max_m |fd_m - g_m| / max_m |g_m|."""
    case = pkg.cases.fourier_case(n, n_steps=steps, checkpoint_every=every)
    s = pkg.Solver(case)
    s.setup_fourier_case()
    s.rk3(2, pkg.MODE_BATCHED, want_dt=False)  # builds the graph
    s.setup_fourier_case()
    s.sync()
    s.rk3(steps, pkg.MODE_BATCHED, want_dt=False)
    s.sync()
    fwd_ms, _ = s.last_step_timing()
    s.setup_fourier_case()
    theta = s.adjoint_forward(steps, every, case["theta"])
    grad = s.adjoint_backward(case["modes"])
    fwd_ckpt_ms, bwd_ms = s.adjoint_timing()
    a = np.array(case["modes"]["amp"])
    fd = np.zeros(len(a))
    for m in range(len(a)):
        th = []
        for sign in (1, -1):
            amp = a.copy()
            amp[m] += sign * h
            s.setup_mixing_layer()
            s.perturb_fourier(0, case["modes"], amp)
            s.initialize()
            s.rk3(steps, pkg.MODE_BATCHED, want_dt=False)
            th.append(s.momentum_thickness(0, case["theta"]))
        fd[m] = (th[0] - th[1]) / (2 * h)
    s.close()
    gmax = float(np.max(np.abs(grad)))
    normwise = float(np.max(np.abs(fd - grad)) / gmax)
    print(f"[bench] adjoint {n}^3 x {steps} steps: forward {fwd_ms:.2f} ms, forward with checkpoints "
          f"{fwd_ckpt_ms:.2f} ms, reverse {bwd_ms:.2f} ms, ratio {bwd_ms / fwd_ms:.2f}; "
          f"|fd - g|_inf / |g|_inf = {normwise:.2e} over {len(a)} amplitudes", file=sys.stderr, flush=True)
    return {"ratio": bwd_ms / fwd_ms, "forward_ms": fwd_ms, "forward_with_checkpoints_ms": fwd_ckpt_ms,
            "reverse_ms": bwd_ms, "theta": theta, "n_params": len(a), "fd_step": h,
            "fd_normwise_rel": normwise,
            "fd_rel_per_mode": [float(abs(fd[m] - grad[m]) / gmax) for m in range(len(a))],
            "workload": f"{n}^3 single-block air mixing layer, WENO-Z5 cp + LF, viscous, {steps} fixed-dt "
                        f"SSP-RK3 steps, checkpoints every {every} steps, 32 Fourier-mode amplitudes -> "
                        f"momentum thickness"}


def characteristic_leg(steps: int):
    """The 24-block air layer with characteristic WENO-Z5 (inviscid_scheme 52):
    autotuned batched steps (graph replay) against the per-block path, same
    state and step count."""
    case = pkg.cases.batched_case(scheme=52)
    cells = case["n_cells"]
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    s.initialize()
    idx, best = s.autotune(2)
    s.rk3(2, pkg.MODE_BATCHED)
    s.sync()
    t0 = time.perf_counter()
    s.rk3(steps, pkg.MODE_BATCHED, want_dt=False)
    s.sync()
    ms_b = (time.perf_counter() - t0) * 1e3 / steps
    s.rk3(1, pkg.MODE_PLAIN)
    s.sync()
    t0 = time.perf_counter()
    s.rk3(steps, pkg.MODE_PLAIN, want_dt=False)
    s.sync()
    ms_p = (time.perf_counter() - t0) * 1e3 / steps
    s.close()
    print(f"[bench] scheme 52, 24 blocks: batched {ms_b:.2f} ms/step (tile {pkg.tile_shapes()[idx]}, "
          f"fused {best:.3f} ms/stage), per-block {ms_p:.2f} ms/step", file=sys.stderr, flush=True)
    return {"workload": "24-block air mixing layer, characteristic WENO-Z5 (inviscid_scheme 52) + LF, viscous",
            "cells": cells, "steps": steps, "value": cells / ms_b / 1e3, "unit": "Mcell-updates/s",
            "ms_per_step": ms_b, "plain_ms_per_step": ms_p, "speedup_vs_plain": ms_p / ms_b,
            "tile": list(pkg.tile_shapes()[idx]), "fused_ms_per_stage": best}


def parallel_leg(reps: int = 10):
    """Pack and unpack of the faces one GPU shares with its neighbours when the
    1024x512x256 reacting case is cut into 8 x-slabs (one 128x512x256 block
    per GPU, faces to the ranks on both sides): 6 primitives + 9 mass
    fractions x (ngg + 1) layers per face node, the reference's message
    (parallel_communication).  Algorithmic HBM bytes per face: pack reads and
    writes the message once; unpack reads the message and the face nodes'
    own 15 values, writes 15 primitives and 14 conservative variables per
    node.  The transfer between GPUs is not measured."""
    MX = pkg.mixture
    tables = MX.load_mechanism(*MX.H2_AIR_FILES)
    nx, ny, nz = 128, 512, 256
    case = pkg.cases.reacting_case([nx - 1], [ny - 1], [nz - 1], tables)
    case["boundaries"] = [b for b in case["boundaries"] if b.get("axis") != 0]
    case["parallel_faces"] = [dict(block=0, range=[1, 1, 1, ny, -1, -nz], rank=0),
                              dict(block=0, range=[nx, nx, 1, ny, -1, -nz], rank=2)]
    s = MX.FrSolver(case, tables)
    s.setup_mixing_layer()
    s.initialize()
    info = s.parallel_faces()
    nv, layers = 6 + s.ns, s.ngg + 1
    L, ctx = s.L, s.ctx

    def pack():
        if L.ostc_fr_pack_parallel(ctx):
            raise RuntimeError(L.ostc_fr_last_error(ctx))

    def pack_unpack():
        pack()
        if L.ostc_fr_unpack_parallel(ctx):
            raise RuntimeError(L.ostc_fr_last_error(ctx))

    out = {}
    for name, fn in (("pack", pack), ("pair", pack_unpack)):
        fn()
        s.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        s.sync()
        out[name + "_ms"] = (time.perf_counter() - t0) * 1e3 / reps
    out["unpack_ms"] = out["pair_ms"] - out["pack_ms"]  # an unpack needs the pack before it
    total = sum(f["count"] for f in info) * 8
    nodes = sum(f["points"] for f in info) * layers
    pack_bytes = 2 * total
    unpack_bytes = total + nv * 8 * (nodes // layers) + total + (nv - 1) * 8 * nodes
    s.close()
    res = {"block": [nx, ny, nz], "faces": len(info), "bytes_per_face": info[0]["count"] * 8,
           "message_bytes": total, "pack_ms": out["pack_ms"], "unpack_ms": out["unpack_ms"],
           "pack_hbm_bytes": pack_bytes, "unpack_hbm_bytes": unpack_bytes,
           "pack_gbs": pack_bytes / out["pack_ms"] / 1e6, "unpack_gbs": unpack_bytes / out["unpack_ms"] / 1e6,
           "pack_frac_hbm": pack_bytes / out["pack_ms"] / 1e6 / HBM_PEAK_GBS,
           "unpack_frac_hbm": unpack_bytes / out["unpack_ms"] / 1e6 / HBM_PEAK_GBS,
           "note": "one GPU's share of the 8-way x-slab cut of the 1024x512x256 reacting case; per call both faces "
                   "(one launch each); algorithmic bytes; the transfer between GPUs (xGMI) is not measured"}
    print(f"[bench] parallel faces of a 128x512x256 reacting slab: {len(info)} faces x "
          f"{info[0]['count'] * 8 / 1e6:.1f} MB, pack {out['pack_ms']:.3f} ms ({res['pack_gbs']:.0f} GB/s), "
          f"unpack {out['unpack_ms']:.3f} ms ({res['unpack_gbs']:.0f} GB/s)", file=sys.stderr, flush=True)
    return res


def reacting_leg(x_cells, y_cells, z_cells, steps: int, warmup: int, label: str, cpu_sample=None, spatial=False):
    """Reacting H2/air mixing layer (9 species, 19 reactions, characteristic
    WENO-Z5 + LF-type splitting, mixture-averaged transport, finite-rate
    chemistry, SSP-RK3 CFL 0.5) through the finite-rate path: `steps`
    graph-replayed steps after `warmup`, then one step with HIP events around
    every kernel for the per-kernel breakdown.  spatial: the spatially
    developing layer (mixing-layer inflow with 1e-3 white noise at x-,
    outflow at x+); otherwise the temporal layer (periodic x)."""
    mx = pkg.mixture
    tables = mx.load_mechanism(*mx.H2_AIR_FILES)
    case = pkg.cases.reacting_case(x_cells, y_cells, z_cells, tables, scheme=52, spatial=spatial)
    cells = case["n_cells"]
    print(f"[bench] {label}: {len(case['blocks'])} blocks, {cells} cells", file=sys.stderr, flush=True)
    s = mx.FrSolver(case, tables)
    s.setup_mixing_layer()
    s.initialize()
    s.rk3(warmup, mx.MODE_BATCHED)
    s.sync()
    t0 = time.perf_counter()
    s.rk3(steps, mx.MODE_BATCHED, want_dt=False)
    s.sync()
    wall = time.perf_counter() - t0
    # the per-block path (one launch per kernel and block, graph-replayed) on the same state
    s.rk3(1, mx.MODE_GRAPH, want_dt=False)
    s.sync()
    t0 = time.perf_counter()
    s.rk3(steps, mx.MODE_GRAPH, want_dt=False)
    s.sync()
    wall_pb = time.perf_counter() - t0
    s.rk3(1, mx.MODE_BATCHED_TIMED, want_dt=False)
    kms = s.kernel_ms()
    dt_last = s.rk3(1, mx.MODE_BATCHED)
    finite = all(bool(np.isfinite(s.get(b, mx.FIELD_BV)).all()) for b in range(len(case["blocks"])))
    s.close()
    value = cells * steps / wall / 1e6
    per_block_value = cells * steps / wall_pb / 1e6
    total_k = sum(v[0] for v in kms.values())
    top = max(kms.items(), key=lambda kv: kv[1][0])
    print(f"[bench] {label}: batched {wall * 1e3 / steps:.2f} ms/step, {value:.1f} Mcell-updates/s; per-block "
          f"{wall_pb * 1e3 / steps:.2f} ms/step; dt {dt_last:.3e}, finite {finite}; kernels (ms/step): " +
          ", ".join(f"{k} {v[0]:.2f}" for k, v in kms.items()), file=sys.stderr, flush=True)
    out = {"value": value, "unit": "Mcell-updates/s", "mode": "BATCHED (one launch per kernel kind over all blocks, "
           "graph-replayed)", "per_block_value": per_block_value, "speedup_vs_per_block": value / per_block_value,
           "ms_per_step": wall * 1e3 / steps, "steps": steps,
           "warmup": warmup, "cells": cells, "blocks": len(case["blocks"]), "state_finite": finite,
           "kernel_ms_per_step": {k: v[0] for k, v in kms.items()},
           "kernel_launches_per_step": {k: int(v[1]) for k, v in kms.items()},
           "dominant_kernel": top[0], "dominant_kernel_share": top[1][0] / total_k,
           "roofline": {"bound": "fp64", "bytes_per_cell_step": BYTES_PER_CELL_STEP_FR,
                        "hbm_frac": value * 1e6 * BYTES_PER_CELL_STEP_FR / (HBM_PEAK_GBS * 1e9),
                        "note": "contract-minimal bytes; the kernels are FP64/transcendental-issue bound"},
           "workload": f"{label}: {'spatially developing (inflow x-, outflow x+)' if spatial else 'temporal'} "
                       f"H2/air mixing layer, 9 species / 19 reactions, WENO-Z5 characteristic "
                       f"+ LF-type splitting, mixture-averaged transport, finite-rate chemistry, SSP-RK3 CFL 0.5"}
    if cpu_sample is not None:
        out["cpu_baseline"] = reacting_cpu_baseline(tables, *cpu_sample, spatial=spatial)
    return out


def reacting_cpu_baseline(tables, n, steps, spatial=False):
    """The CPU oracle of the finite-rate path on a bounded sample: `steps`
    RK3 steps of an n x n/2 x n/2 block of the same reacting layer."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mech
    import oracle_fr as OF

    threads = cpu_threads()
    to = mech.load(*pkg.mixture.H2_AIR_FILES)
    case = pkg.cases.reacting_single_block(n, n // 2, n // 2, tables, scheme=52, spatial=spatial)
    o = OF.FrOracle(case, to, n_threads=threads)
    o.setup_mixing_layer()
    o.initialize()
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    cells = case["n_cells"]
    return {"value": cells * steps / dt / 1e6, "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"{steps} RK3 steps of a {n}x{n // 2}x{n // 2} block ({cells} cells) of the same reacting "
                      f"layer, {dt:.1f} s on {threads} OpenMP threads"}


def measured_traffic(tile, split: int, waves: int):
    """HBM bytes per launch of the fused kernel for this tile configuration,
    from the FETCH_SIZE / WRITE_SIZE counter runs summarised in
    profiles/fused_rhs_update_batched24_pmc.txt (the same 24-block case), or
    None when the configuration was not measured."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "fused_rhs_update_batched24_pmc.txt")
    if waves != 0 or not os.path.exists(path):
        return None
    want = f"traffic tile={tile[0]}x{tile[1]}x{tile[2]} split={int(split)} "
    with open(path) as f:
        for line in f:
            if line.startswith(want):
                fields = dict(kv.split("=") for kv in line.split()[1:])
                return int(fields["bytes"])
    return None


def flop_counts(path=os.path.join(ROOT, "profiles", "flops_per_cell_air.txt")):
    """FP64 operations per interior cell counted by the oracle's counting
    build (tests/golden/make_flop_counts.py): {name: flops} of each line."""
    out = {}
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            name, *fields = line.split()
            out[name] = float(dict(kv.split("=") for kv in fields)["flops"])
    return out


def fused_roofline(cells, avg_kernel_ms, rate_mcells, tile_index, tile_split, tile_waves):
    """Roofline of the fused RHS+update kernel (one RK stage over all blocks)
    and of the whole step.  The kernel is measured against both roofs: HBM
    with the contract-minimal 264 B per cell and stage, FP64 with the
    oracle-counted flops per cell and stage; `bound` is the roof that binds
    (the lower attainable rate), `achieved`/`peak`/`frac` are on that roof.
    step: rate / min(8e12 / B_step, P_FP64 / F_step) in cell-updates/s."""
    f = flop_counts()
    avg_s = avg_kernel_ms / 1e3
    bytes_per_launch = BYTES_PER_CELL_STAGE * cells
    flops_per_launch = f["kernel_stage_per_cell"] * cells
    gbs = bytes_per_launch / avg_s / 1e9
    tfs = flops_per_launch / avg_s / 1e12
    # attainable launches/s on each roof; the smaller one binds
    hbm_bound = HBM_PEAK_GBS * 1e9 / bytes_per_launch < FP64_PEAK_TFLOPS * 1e12 / flops_per_launch
    b_step = 3 * BYTES_PER_CELL_STAGE + 8 * 5
    step_roof = min(HBM_PEAK_GBS * 1e9 / b_step, FP64_PEAK_TFLOPS * 1e12 / f["step_per_cell"]) / 1e6
    traffic = measured_traffic(pkg.tile_shapes()[tile_index], tile_split, tile_waves)
    hbm = {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}
    fp64 = {"achieved": tfs, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tfs / FP64_PEAK_TFLOPS}
    main = hbm if hbm_bound else fp64
    return {"bound": "hbm" if hbm_bound else "fp64", **main, "traffic": traffic,
            "traffic_source": "profiles/fused_rhs_update_batched24_pmc.txt (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                              "copy-calibrated, bytes per launch)" if traffic else None,
            "kernel": "k_fused (RHS+update, one RK stage over all blocks)",
            "hbm": hbm, "fp64": fp64,
            "algorithmic_bytes_per_launch": bytes_per_launch, "algorithmic_flops_per_launch": flops_per_launch,
            "flops_source": "profiles/flops_per_cell_air.txt (oracle on a counting scalar; fp64 vector peak)",
            "avg_launch_ms": avg_kernel_ms,
            "step": {"bytes_per_cell": b_step, "flops_per_cell": f["step_per_cell"],
                     "roof_mcell_updates_per_s": step_roof, "frac": rate_mcells / step_roof,
                     "hbm_frac": rate_mcells * 1e6 * b_step / (HBM_PEAK_GBS * 1e9),
                     "fp64_frac": rate_mcells * 1e6 * f["step_per_cell"] / (FP64_PEAK_TFLOPS * 1e12)}}


def copy_probe(mib: int):
    """Five launches of an elementwise FP64 copy (b = a * 1) of ``mib`` MiB:
    each reads and writes exactly mib MiB, the reference figure that turns
    the FETCH_SIZE / WRITE_SIZE counters of a profiled run into bytes."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read_per_launch": n * 8, "bytes_written_per_launch": n * 8,
                      "launches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--tile", type=int, default=-1, metavar="INDEX",
                    help="use this entry of the tile table (one workgroup per tile) instead of the autotuner")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-characteristic", action="store_true", help="skip the scheme-52 batched/per-block leg")
    ap.add_argument("--adjoint-n", type=int, default=128)
    ap.add_argument("--adjoint-steps", type=int, default=50)
    ap.add_argument("--adjoint-every", type=int, default=10)
    ap.add_argument("--no-reacting", action="store_true")
    ap.add_argument("--reacting-steps", type=int, default=5)
    ap.add_argument("--no-large", action="store_true", help="skip the 134M-cell reacting case")
    ap.add_argument("--reacting-only", action="store_true",
                    help="run only the 512x256x128 reacting case and print its result (profiling)")
    ap.add_argument("--parallel-only", action="store_true",
                    help="only the pack/unpack leg of the parallel faces (one JSON line)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain FP64 copy of MIB MiB five times (calibrates HBM byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: this build runs the path on a "
              "single GPU and does not shard it.", file=sys.stderr)
        sys.exit(2)

    if a.parallel_only:
        print(json.dumps(parallel_leg(reps=20)), flush=True)
        return
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return

    if a.reacting_only:
        r = reacting_leg([511], [255], [127], a.reacting_steps, 2, "512x256x128 single block", spatial=True)
        print(json.dumps({"reacting": {"single_block_512x256x128": r}}))
        return

    case = pkg.cases.batched_case()
    cells = case["n_cells"]
    print(f"[bench] {len(case['blocks'])} blocks, {cells} interior cells", file=sys.stderr, flush=True)
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    s.initialize()
    if a.tile >= 0:
        s.set_tile(a.tile, 0, 0)
    elif not a.no_tune:
        best = s.autotune(3)
        print(f"[bench] autotune: best {best[1]:.3f} ms/stage", file=sys.stderr, flush=True)
    tile_index, tile_waves, tile_split = s.tile_config()
    print(f"[bench] fused kernel config: tile {pkg.tile_shapes()[tile_index]}, "
          f"{tile_waves} workgroups/CU (0 = one per tile), split={tile_split}", file=sys.stderr, flush=True)

    # BATCHED: warm-up, then exactly K steps.  The timed region launches the
    # batched kernels directly with HIP events around every fused RHS+update
    # launch (the kernel durations behind the roofline figure); the same
    # steps replayed as a captured graph are timed afterwards as value_graph.
    s.rk3(a.warmup, pkg.MODE_BATCHED)
    s.sync()
    t0 = time.perf_counter()
    s.rk3(a.steps, pkg.MODE_BATCHED_TIMED, want_dt=False)
    s.sync()
    wall = time.perf_counter() - t0
    dev_ms, launches = s.last_step_timing()
    avg_kernel_ms, max_kernel_ms, n_kernel = s.fused_kernel_ms()
    value = cells * a.steps / wall / 1e6
    print(f"[bench] batched: {wall * 1e3 / a.steps:.3f} ms/step wall, {dev_ms / a.steps:.3f} ms/step device, "
          f"fused kernel avg {avg_kernel_ms:.3f} ms max {max_kernel_ms:.3f} ms over {n_kernel} launches, "
          f"{launches // max(a.steps, 1)} launches/step", file=sys.stderr, flush=True)
    s.sync()
    t0 = time.perf_counter()
    s.rk3(a.steps, pkg.MODE_BATCHED, want_dt=False)
    s.sync()
    wall_g = time.perf_counter() - t0
    value_graph = cells * a.steps / wall_g / 1e6
    print(f"[bench] batched (graph replay): {wall_g * 1e3 / a.steps:.3f} ms/step", file=sys.stderr, flush=True)

    # PLAIN (reference-shaped per-block launches, host dt read-back each step)
    s.rk3(2, pkg.MODE_PLAIN)
    s.sync()
    t0 = time.perf_counter()
    s.rk3(a.plain_steps, pkg.MODE_PLAIN, want_dt=False)
    s.sync()
    wall_p = time.perf_counter() - t0
    _, launches_p = s.last_step_timing()
    plain_value = cells * a.plain_steps / wall_p / 1e6
    print(f"[bench] plain: {wall_p * 1e3 / a.plain_steps:.3f} ms/step, "
          f"{launches_p // max(a.plain_steps, 1)} launches/step", file=sys.stderr, flush=True)

    roofline = fused_roofline(cells, avg_kernel_ms, value, tile_index, tile_split, tile_waves)
    s.close()

    characteristic = None if a.no_characteristic else characteristic_leg(5)
    parallel = None if (a.no_reacting or a.no_large) else parallel_leg()

    adjoint = None
    if not a.no_adjoint:
        adjoint = adjoint_leg(a.adjoint_n, a.adjoint_steps, a.adjoint_every)

    reacting = None
    if not a.no_reacting:
        reacting = {"single_block_512x256x128": reacting_leg(
            [511], [255], [127], a.reacting_steps, 2, "512x256x128 single block",
            cpu_sample=None if a.no_cpu else (96, 2), spatial=True)}
        if not a.no_large:
            rng = np.random.default_rng(42)
            xs = [int(v) for v in rng.integers(384, 640, size=1)]
            ys = [int(v) for v in rng.integers(192, 320, size=1)]
            zs = [int(v) for v in rng.integers(96, 160, size=1)]
            reacting["large_1024x512x256_8_blocks"] = reacting_leg(
                [xs[0], 1023 - xs[0]], [ys[0], 511 - ys[0]], [zs[0], 255 - zs[0]], a.reacting_steps, 1,
                "1024x512x256 in 8 uneven blocks")

    cpu = None
    if not a.no_cpu:
        print(f"[bench] cpu baseline: {a.cpu_steps} oracle steps", file=sys.stderr, flush=True)
        cpu = cpu_baseline(case, a.cpu_steps)

    out = {
        "metric": METRIC, "value": value, "unit": "Mcell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": wall * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batched multi-block air temporal mixing layer, 24 uneven blocks (cuts of 16..192 x "
                               f"16..128 x 16..64 cells, seed {pkg.cases.batched_seed()}), "
                               "WENO-Z5 cp + LF, 2nd-order viscous, SSP-RK3 CFL 0.5",
                   "cells": cells, "blocks": len(case["blocks"]), "mode": "BATCHED",
                   "tile": list(pkg.tile_shapes()[tile_index]), "tile_index": tile_index,
                   "workgroups_per_cu": tile_waves,
                   "split_stage": bool(tile_split)},
        "value_graph": value_graph, "plain_value": plain_value, "speedup_vs_plain": value / plain_value,
        "roofline": roofline, "cpu_baseline": cpu, "characteristic": characteristic, "adjoint": adjoint,
        "reacting": reacting, "parallel_faces": parallel,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
