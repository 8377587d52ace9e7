"""The C-ABI library loads on a machine without a GPU and exports every entry
point include/lbm_amd.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbm_amd.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from lbm import _native as nat
    names = declared()
    assert "lbm_step" in names and "lbm_create" in names
    missing = [n for n in names if not hasattr(nat.lib, n)]
    assert not missing, missing


def test_params_default_and_error_string_without_gpu():
    from lbm import _native as nat
    p = nat.LbmParams()
    nat.lib.lbm_params_default(ctypes.byref(p), 16, 8, 4)
    assert (p.nx, p.ny, p.nz) == (16, 8, 4)
    assert p.boundary_relax == 0.3 and p.sharpen_reduce == 0.8
    assert p.precision == nat.LBM_F64
    # a refused argument reports through lbm_last_error without touching the GPU
    rc = nat.lib.lbm_create(None, None)
    assert rc == -1
    assert b"null" in nat.lib.lbm_last_error()


def test_params_struct_layout_matches_header():
    """ctypes mirror of lbm_params has the field order of the header."""
    from lbm import _native as nat
    src = open(HEADER).read()
    body = src[src.index("typedef struct lbm_params"):src.index("} lbm_params;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"(?:int|double)\s+([^;]+);", body)
    names = []
    for f in fields:
        for n in f.split(","):
            names.append(n.strip().split("[")[0])
    assert names == [f[0] for f in nat.LbmParams._fields_]
