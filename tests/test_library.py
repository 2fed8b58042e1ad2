"""The C-ABI library loads and exports every entry point include/mdmc.h declares (no GPU needed)."""
import ctypes
import os
import re

import manifolddmc_amd as mdmc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "mdmc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mdmc_[a-z_]+)\s*\(", src)))


def test_header_declares_the_python_exports():
    assert declared_functions() == sorted(mdmc.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(mdmc.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_options_struct_layout_matches_header():
    assert ctypes.sizeof(mdmc._Options) == 16
    assert ctypes.sizeof(mdmc._Grid) == 3 * 4 * 3 + 4 + 8 + 4 + 4  # dims, origin, delta, pad, ptr, flag, tail pad
