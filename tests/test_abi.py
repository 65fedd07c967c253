"""The C-ABI library loads and exports every function include/recblock.h
declares (no compute: runs without a GPU)."""
import ctypes
import os
import re

import rbsptrsv
from rbsptrsv import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "recblock.h")


def _declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(rb_\w+|recblocking_\w+)\s*\(", src)
    return sorted(set(names))


def test_header_and_export_list_agree():
    assert sorted(_lib.EXPORTS) == _declared_functions()


def test_library_exports_every_symbol():
    assert os.path.exists(rbsptrsv.LIB_PATH), "librecblock.so not built"
    L = ctypes.CDLL(rbsptrsv.LIB_PATH)
    for name in _declared_functions():
        assert hasattr(L, name), name


def test_host_only_entry_points():
    L = _lib.lib()
    assert L.rb_lv_auto(2_250_000, 0) == 4
    assert L.rb_lv_auto(4_000_000, 0) == 5
    assert L.rb_lv_auto(100_000, 0) == 1
    old = L.rb_set_bench_repeat(7)
    assert L.rb_set_bench_repeat(old) == 7
    # argument errors are reported, never aborted on, and need no device
    h = ctypes.c_void_p()
    st = L.rb_plan_create(None, None, None, 4, 5, 4, 0, 1, None, None,
                          ctypes.byref(h))
    assert st == 2  # RB_ERR_NOT_SQUARE
    assert b"m != n" in L.rb_last_error()
    assert L.rb_solve(None, None, 1, None, 1, 1, None) == 1
