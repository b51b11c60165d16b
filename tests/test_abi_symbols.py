"""The C-ABI library loads on a machine without a GPU and exports every
function that include/openstc/*.h declares (no compute calls)."""
import ctypes
import glob
import os
import re

from _util import ROOT, pkg


def declared_functions():
    text = "".join(open(p).read() for p in sorted(glob.glob(os.path.join(ROOT, "include", "openstc", "*.h"))))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ostc_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_operator_surface():
    names = declared_functions()
    for must in ("ostc_create", "ostc_step", "ostc_stage_rhs", "ostc_compute_dt", "ostc_stage_update",
                 "ostc_apply_bc", "ostc_exchange", "ostc_update_properties", "ostc_residual"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(pkg.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_binds_every_symbol():
    assert set(declared_functions()) == set(pkg.SIGNATURES)


def test_tile_table_is_queryable_without_gpu():
    shapes = pkg.tile_shapes()
    assert len(shapes) >= 4
    for tx, ty, tz in shapes:
        assert tx * ty * tz in (128, 256, 512)
