"""Caller-owned device arrays at the block boundary (needs a GPU).

A caller that keeps its own DZone arrays (Field::setup_device_memory,
src/Field.cu:469-647) passes their device addresses in ostc_block_desc; the
library adopts them instead of allocating its own, and reads a device metric
in the reference's 80-byte gxl::Matrix AoS (element stride 10, component
stride 1, offset 1; src/gxl_lib/Matrix.hpp:53-72) where it lies.  The same
steps on library-owned and on caller-owned arrays must give bit-identical
states, and the caller's arrays must hold the library's state afterwards.
The device arrays are allocated with the HIP runtime the library links.
"""
import ctypes as C

import numpy as np
import pytest

from _util import cases, pkg
import mech

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime the library links."""

    def __init__(self):
        pkg.lib()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []

    def alloc(self, n_doubles):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), n_doubles * 8) == 0
        assert self.L.hipMemcpy(p, np.zeros(n_doubles).ctypes.data, n_doubles * 8, H2D) == 0
        self.ptrs.append(p)
        return p.value

    def upload(self, ptr, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert self.L.hipMemcpy(ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def download(self, ptr, shape):
        a = np.empty(shape)
        assert self.L.hipMemcpy(a.ctypes.data, ptr, a.nbytes, D2H) == 0
        return a

    def free(self):
        for p in self.ptrs:
            self.L.hipFree(p)
        self.ptrs = []


def _aos_metric(met):
    """SoA metric planes [9][...] -> the 10-double-per-point AoS with the
    components at offsets 1..9 (slot 0 holds a marker the library must skip)."""
    n = met[0].size
    aos = np.full((n, 10), -7.0)
    aos[:, 1:] = met.reshape(9, n).T
    return aos


def _arrays(hip, shp, nv, n0, ns=0):
    sz = int(np.prod(shp))
    arr = dict(cv=hip.alloc(nv * sz), bv=hip.alloc(6 * sz), qn=hip.alloc(nv * sz), dq=hip.alloc(nv * n0),
               mul=hip.alloc(sz), jac=hip.alloc(sz), metric=hip.alloc(10 * sz), metric_strides=(10, 1, 1))
    if ns:
        arr["sv"] = hip.alloc(ns * sz)
    return arr


def test_air_blocks_on_caller_arrays():
    case = cases.small_multiblock_case()
    print(f"[gpu] caller-owned arrays, air, {len(case['blocks'])} blocks", flush=True)
    ref = pkg.Solver(case)
    ref.setup_mixing_layer()
    g = ref.ngg
    hip = Hip()
    try:
        arrays = []
        for b, blk in enumerate(case["blocks"]):
            shp = pkg.padded_shape(blk, g)
            a = _arrays(hip, shp, 5, blk["mx"] * blk["my"] * blk["mz"])
            hip.upload(a["metric"], _aos_metric(ref.get(b, pkg.FIELD_METRIC)))
            hip.upload(a["jac"], ref.get(b, pkg.FIELD_JAC))
            arrays.append(a)
        s = pkg.Solver(case, device_arrays=arrays)
        for b in range(len(case["blocks"])):
            np.testing.assert_array_equal(s.get(b, pkg.FIELD_METRIC), ref.get(b, pkg.FIELD_METRIC))
            np.testing.assert_array_equal(s.get(b, pkg.FIELD_JAC), ref.get(b, pkg.FIELD_JAC))
            s.set(b, pkg.FIELD_BV, ref.get(b, pkg.FIELD_BV))
        for x in (ref, s):
            x.initialize()
            x.rk3(3, pkg.MODE_BATCHED)
            x.rk3(2, pkg.MODE_PLAIN)
        for b, blk in enumerate(case["blocks"]):
            shp = pkg.padded_shape(blk, g)
            for f, name, nplanes in ((pkg.FIELD_BV, "bv", 6), (pkg.FIELD_CV, "cv", 5), (pkg.FIELD_MUL, "mul", 1)):
                want = ref.get(b, f)
                np.testing.assert_array_equal(s.get(b, f), want)
                np.testing.assert_array_equal(hip.download(arrays[b][name], (nplanes, *shp)), want)
        s.close()
        ref.close()
    finally:
        hip.free()


def test_mixture_block_on_caller_arrays():
    MX = pkg.mixture
    tg = MX.load_mechanism(*MX.H2_AIR_FILES)
    case = cases.reacting_single_block(12, 14, 10, tg, scheme=52)
    print("[gpu] caller-owned arrays, mixture", flush=True)
    ref = MX.FrSolver(case, tg)
    ref.setup_mixing_layer()
    g, ns, nv = ref.ngg, ref.ns, ref.nv
    blk = case["blocks"][0]
    shp = pkg.padded_shape(blk, g)
    hip = Hip()
    try:
        a = _arrays(hip, shp, nv, blk["mx"] * blk["my"] * blk["mz"], ns)
        hip.upload(a["metric"], _aos_metric(ref.get(0, MX.FIELD_METRIC)))
        hip.upload(a["jac"], ref.get(0, MX.FIELD_JAC))
        s = MX.FrSolver(case, tg, device_arrays=[a])
        np.testing.assert_array_equal(s.get(0, MX.FIELD_METRIC), ref.get(0, MX.FIELD_METRIC))
        s.set(0, MX.FIELD_BV, ref.get(0, MX.FIELD_BV))
        s.set(0, MX.FIELD_SV, ref.get(0, MX.FIELD_SV))
        for x in (ref, s):
            x.initialize()
            x.rk3(3, MX.MODE_GRAPH)
        for f, name, nplanes in ((MX.FIELD_BV, "bv", 6), (MX.FIELD_SV, "sv", ns), (MX.FIELD_CV, "cv", nv)):
            want = ref.get(0, f)
            np.testing.assert_array_equal(s.get(0, f), want)
            np.testing.assert_array_equal(hip.download(a[name], (nplanes, *shp)), want)
        s.close()
        ref.close()
    finally:
        hip.free()
