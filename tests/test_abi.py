"""The C ABI library loads without a GPU and exports every entry point that
include/mhd2d.h declares; the binding's struct matches the header's."""
import ctypes
import os
import re

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
