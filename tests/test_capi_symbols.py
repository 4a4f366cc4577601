"""The C-ABI library loads without a GPU and exports every entry point declared
in include/*.h; the ctypes binding declares exactly those entry points."""
import ctypes
import glob
import os
import re

from conftest import ROOT


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(spr_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert {"spr_create", "spr_destroy", "spr_iterate", "spr_set_population", "spr_get_population"} <= names


def test_library_exports_all_declared_symbols():
    from s_spi_spring_amd import hipcore
    assert os.path.exists(hipcore.LIB_PATH), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(hipcore.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from s_spi_spring_amd import hipcore
    assert set(hipcore.SIGNATURES) == declared_functions()


def test_load_and_size_query_without_gpu():
    from s_spi_spring_amd import hipcore
    lib = hipcore.load()
    assert lib.spr_supported_size(512) == 1
    assert lib.spr_supported_size(500) == 0
    # spr_create reports the failure instead of crashing when no device is present
    import numpy as np
    amp = np.zeros((64, 64), np.float32)
    h = ctypes.c_void_p()
    rc = lib.spr_create(ctypes.byref(h), 64, 1, amp.ctypes.data_as(ctypes.c_void_p), 99)
    assert rc != 0
    assert lib.spr_last_error(None)
