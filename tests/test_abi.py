"""The C ABI library loads without a GPU and exports every entry point that
include/mhd2d.h declares; the binding's struct matches the header's."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import PKG_DIR, ROOT

HEADER = os.path.join(ROOT, "include", "mhd2d.h")
LIB = os.path.join(PKG_DIR, "libmhd2d.so")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mhd2d_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("mhd2d_create", "mhd2d_lh", "mhd2d_lhez", "mhd2d_tvb_limiter_p2", "mhd2d_pp_limiter",
                 "mhd2d_kxrcf_detector", "mhd2d_div_free_balsara", "mhd2d_calculate_dt", "mhd2d_step"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmhd2d.so not built")
def test_library_exports_every_declared_symbol(pkg):
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the binding declares exactly the header's functions
    assert sorted(pkg.EXPORTS) == declared_functions()


def test_config_struct_layout(pkg):
    # mhd2d_config: 12 int32 then 9 doubles, no padding
    assert ctypes.sizeof(pkg.Config) == 12 * 4 + 9 * 8
    text = open(HEADER).read()
    body = text[text.index("typedef struct mhd2d_config"):text.index("} mhd2d_config;")]
    fields = re.findall(r"(int32_t|double)\s+([\w, ]+);", body)
    names = [n.strip() for _, group in fields for n in group.split(",")]
    assert names == [f[0] for f in pkg.Config._fields_]


def test_info_struct_layout(pkg):
    assert ctypes.sizeof(pkg.ICInfo) == 6 * 8 + 6 * 4
    text = open(HEADER).read()
    body = text[text.index("typedef struct mhd2d_ic_info"):text.index("} mhd2d_ic_info;")]
    fields = re.findall(r"(int32_t|double)\s+([\w, ]+);", body)
    names = [n.strip() for _, group in fields for n in group.split(",")]
    assert names == [f[0] for f in pkg.ICInfo._fields_]


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmhd2d.so not built")
@pytest.mark.parametrize("tag,ic", [("vortex", 1), ("ot", 3), ("rotor", 4), ("blast", 5), ("cloud", 6),
                                    ("cloud2", 61), ("loop", 7), ("briowu", 8), ("briowu_y", 81),
                                    ("ryujones", 9), ("ryujones_y", 91)])
def test_ic_defaults_match_reference(pkg, tag, ic):
    """mhd2d_ic_defaults (host only, no device) returns what the reference's
    init_data leaves in its COMMON blocks for that initN.txt"""
    from conftest import golden
    g = golden(f"init_{tag}_16")
    d = pkg.ic_defaults(ic)
    assert d["domain"] == tuple(float(v) for v in g["domain"])
    assert d["M"] == float(g["M"]) and d["tend"] == float(g["tend"])
    if "bc" in g:
        assert d["bc"] == (int(g["bc"]),) * 4
    with pytest.raises(pkg.MHD2DError):
        pkg.ic_defaults(12345)


def test_fortran_binding_covers_every_export():
    """fortran/mhd2d_capi.f90 has one bind(C) interface per function the header
    declares, under the same C name"""
    f = open(os.path.join(PKG_DIR, "fortran", "mhd2d_capi.f90")).read()
    bound = sorted(set(re.findall(r"name='(mhd2d_\w+)'", f)))
    assert bound == declared_functions()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmhd2d.so not built")
def test_calculate_ez_closed_form(pkg):
    """calculate_Ez.f90:5, Ez = u2*B1 - u1*B2: zero for u parallel to B, odd
    in the velocity, and the reference's product order (host entry point, no
    device needed)"""
    rng = np.random.default_rng(5)
    u1, u2, B1, B2 = rng.standard_normal((4, 257))
    ez = pkg.calculate_Ez(u1, u2, B1, B2)
    assert np.array_equal(ez, u2 * B1 - u1 * B2)
    assert np.array_equal(pkg.calculate_Ez(-u1, -u2, B1, B2), -ez)
    assert np.all(pkg.calculate_Ez(B1, B2, B1, B2) == 0.0)
    lib = ctypes.CDLL(LIB)
    assert lib.mhd2d_calculate_ez(ctypes.c_int64(-1), None, None, None, None, None) == -1
