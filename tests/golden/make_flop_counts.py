"""Floating-point operations per cell of the air RHS + SSP-RK3 path, counted
by the oracle's operation-counting build (-DORC_COUNT; the counts of its
routines are pinned in tests/test_oracle_flop_count.py).

One flop = one addition/subtraction, multiplication, division or square
root; calls of pow/tanh/exp are listed apart ("trans").  The workload is a
single block with the mean block shape of the 24-block batched case
(cuts of cases.batched_case(): 140 x 70 x 51 cells, i.e. 141 x 71 x 52
nodes), periodic x/z, outflow y, WENO-Z5 + LF, viscous: its per-cell counts
carry the same share of block-edge faces as the batched case.

Writes profiles/flops_per_cell_air.txt.  Usage:
    python tests/golden/make_flop_counts.py
"""
from __future__ import annotations

import importlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
pkg = importlib.import_module("openstc-mixinglayer_amd")
import oracle as O  # noqa: E402

OUT = os.path.join(ROOT, "profiles", "flops_per_cell_air.txt")


def flops(c):
    return c["add"] + c["mul"] + c["div"] + c["sqrt"]


def fmt(c, n):
    return " ".join(f"{k}={c[k] / n:.2f}" for k in O.OP_NAMES) + f" flops={flops(c) / n:.2f}"


def main():
    xs, ys, zs = pkg.cases._batched_cuts(pkg.cases.batched_seed())
    nx, ny, nz = round(np.mean(xs)), round(np.mean(ys)), round(np.mean(zs))
    case = pkg.cases.mixing_layer_case([nx], [ny], [nz])
    o = O.Oracle(case, tangent="count")
    o.setup_mixing_layer()
    o.initialize()
    L, g = o.L, o.ngg
    blk = case["blocks"][0]
    cells = blk["mx"] * blk["my"] * blk["mz"]
    padded = (blk["mx"] + 2 * g) * (blk["my"] + 2 * g) * (blk["mz"] + 2 * g)
    parts = {}

    def count(name, fn):
        O.op_counts_reset(o)
        fn()
        parts[name] = O.op_counts(o)

    count("step", lambda: o.step(1))
    count("inviscid", lambda: L.orc_inviscid(o.ctx, 0))
    count("viscous", lambda: L.orc_viscous(o.ctx, 0))
    count("update", lambda: L.orc_update(o.ctx, 0, 1e-9, 1))
    count("properties", lambda: L.orc_properties(o.ctx, 0))
    count("local_dt", lambda: L.orc_local_dt(o.ctx, 0))
    # the batched kernel's work per interior cell and stage: inviscid, viscous,
    # the update, and the properties of the interior points
    kern = {k: parts["inviscid"][k] + parts["viscous"][k] + parts["update"][k] +
            parts["properties"][k] * cells / padded for k in O.OP_NAMES}
    lines = [
        "# FP64 operations per interior cell of the air RHS + SSP-RK3 path (WENO-Z5 cp + LF, 2nd-order viscous),",
        "# counted by the oracle built on a counting scalar (-DORC_COUNT, oracle/count.h).  One flop = one",
        "# add/sub, mul, div or sqrt; trans = calls of pow/tanh/exp, not in flops.",
        "# command: python tests/golden/make_flop_counts.py",
        f"# workload: one block of {blk['mx']}x{blk['my']}x{blk['mz']} nodes ({cells} cells; the mean block of the "
        f"24-block batched case), periodic x/z, outflow y",
        f"step_per_cell {fmt(parts['step'], cells)}",
        f"kernel_stage_per_cell {fmt(kern, cells)}",
    ]
    for name in ("inviscid", "viscous", "update", "local_dt"):
        lines.append(f"{name}_per_cell {fmt(parts[name], cells)}")
    lines.append(f"properties_per_padded_point {fmt(parts['properties'], padded)}")
    text = "\n".join(lines) + "\n"
    with open(OUT, "w") as f:
        f.write(text)
    print(text, end="")


if __name__ == "__main__":
    main()
