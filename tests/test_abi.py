"""The C-ABI library loads and exports exactly what include/tsg.h declares
(no GPU calls)."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

import tilespgemm_amd as tsg

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "tsg.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n.startswith(("tsg_", "thmkl_"))))


def test_header_declares_the_reference_entry():
    assert "thmkl_dcsrmultcsr" in declared_functions()


@pytest.mark.skipif(not os.path.exists(tsg.lib_path()), reason="libtsg.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(tsg.lib_path())
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert sorted(tsg.SIGNATURES) == declared_functions()


@pytest.mark.skipif(not os.path.exists(tsg.lib_path()), reason="libtsg.so not built")
def test_binding_loads_and_reports_errors_without_gpu_work():
    lib = tsg.load()
    # invalid arguments are rejected before any device work
    assert lib.tsg_set_method(None, 1) == -1
    assert lib.tsg_destroy(None) == 0
    assert lib.tsg_workspace_bytes(None) == 0


@pytest.mark.skipif(not os.path.exists(tsg.lib_path()), reason="libtsg.so not built")
def test_host_slab_setting_without_gpu_work():
    """The host entries' slab cap is a plain setting: negative caps are
    rejected, others read back; no call has run, so no slabs either."""
    lib = tsg.load()
    assert lib.tsg_host_last_slabs() == 0
    assert lib.tsg_set_host_slab_entries(-1) == -1
    try:
        tsg.set_host_slab_entries(12345)
        assert tsg.host_slab_entries() == 12345
        with pytest.raises(tsg.TsgError):
            tsg.set_host_slab_entries(-5)
        assert tsg.host_slab_entries() == 12345
    finally:
        tsg.set_host_slab_entries(0)
    assert tsg.host_slab_entries() == 0


def test_plan_array_enum_matches_header():
    text = open(HEADER).read()
    block = text[text.index("TSG_A_TILE_PTR = 0"):text.index("TSG_PLAN_ARRAY_COUNT")]
    names = re.findall(r"TSG_([A-Z_]+?)\s*(?:=\s*0)?,", block)
    assert names == list(tsg.PLAN_ARRAYS)
