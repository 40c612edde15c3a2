"""The C-ABI library loads and exports every entry point include/hss_hip.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import pathlib
import re

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "hss_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hss_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_drop_in_entry_points():
    names = declared_functions()
    for must in ("hss_matmul", "hss_matmul_dev", "hss_ulvfactsolve", "hss_ulv_factor", "hss_ulv_solve",
                 "hss_mat_create", "hss_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol(hip):
    lib = hip.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(hip.EXPORTS) == declared_functions()


def test_node_record_layout_matches_header():
    from conftest import H
    # hss_node_desc: 10 int32 + 9 int64 = 112 bytes, same field order as NODE_DTYPE
    assert H.NODE_DTYPE.itemsize == 112
    text = HEADER.read_text()
    assert "} hss_node_desc; /* 112 bytes */" in text
    names = list(H.NODE_DTYPE.names)
    assert names[:10] == ["parent", "left", "right", "depth", "height", "m", "n", "ku", "kv", "pad"]


def test_oracle_node_record_matches():
    src = (ROOT / "oracle" / "hss_oracle.c").read_text()
    assert "int32_t parent, left, right, depth, height, m, n, ku, kv, pad;" in src
    assert "int64_t row_off, col_off, off_D, off_U, off_V, off_B12, off_B21, off_R, off_W;" in src


def test_refine_opts_modes(hip):
    """Single solves map tol / max_iter / method onto hss_refine_opts: mode 1 is iterative
    refinement, mode 2 GMRES-IR, mode 0 when tol <= 0; no options without tol or max_iter."""
    assert hip._refine_opts(None, None, "ir") is None
    o = hip._refine_opts(1e-10, 5, "ir")
    assert (o.tol, o.max_iter, o.mode) == (1e-10, 5, 1)
    o = hip._refine_opts(1e-9, None, "gmres")
    assert (o.tol, o.max_iter, o.mode) == (1e-9, 10, 2)
    assert hip._refine_opts(None, 3, "gmres").mode == 0
    import pytest
    with pytest.raises(ValueError):
        hip._refine_opts(1e-10, 5, "cg")
