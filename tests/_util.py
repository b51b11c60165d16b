"""Shared helpers for the tests (imports the package and the oracle)."""
import importlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

pkg = importlib.import_module("openstc-mixinglayer_amd")
cases = pkg.cases
import oracle as O  # noqa: E402


def rel_err(a, b, axis=(1, 2, 3)):
    """max|a-b| / max|b| per leading-axis variable."""
    den = np.max(np.abs(b), axis=axis)
    den = np.where(den == 0, 1.0, den)
    return np.max(np.abs(a - b), axis=axis) / den


def interior(arr, g):
    return arr[(slice(None), slice(g, -g), slice(g, -g), slice(g, -g))]


def oracle_for(case, steps_init=True):
    o = O.Oracle(case)
    o.setup_mixing_layer()
    if steps_init:
        o.initialize()
    return o


def oracle_pair(pair, bv_global, tangent=False):
    """The oracle on the rotated two-block case: node coordinates of each
    block's frame, metrics from them (Block::compute_jac_metric), the initial
    primitives mapped from a global single-block padded array, then
    initialize (boundary conditions, exchange, properties, cv)."""
    o = O.Oracle(pair, tangent=tangent)
    g = o.ngg
    L = O.lib(tangent)
    for b in range(len(pair["blocks"])):
        for which, arr in zip((O.X, O.Y, O.Z), cases.block_coordinates(pair, b, g)):
            o.array(b, which)[0][...] = arr
        assert L.orc_compute_metric(o.ctx, b) == 0
        o.array(b, O.BV)[...] = cases.block_from_global(pair, b, bv_global, g)
    o.initialize()
    return o
