"""CPU-side checks of the product library: it loads and exports every entry point of
include/prt_hip.h (no compute calls without a GPU)."""
import ctypes
import os
import re

import protoray_hip as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "prt_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(prt_[a-z0-9_]+)\s*\(", text, re.M)))


def test_header_and_binding_agree():
    assert declared_entry_points() == sorted(P.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = P.load_library()
    for name in declared_entry_points():
        assert hasattr(lib, name), name


def test_last_error_is_callable_without_a_device():
    lib = P.load_library()
    assert isinstance(lib.prt_last_error(), bytes)
    # a null context is rejected with PRT_ERR_INVALID, no GPU needed
    assert lib.prt_frame_clear(None) == P.PRT_ERR_INVALID
    assert b"null context" in lib.prt_last_error()
