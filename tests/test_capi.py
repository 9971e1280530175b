"""The C-ABI library loads and exports every entry point include/rfd.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

from robotfd_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "rfd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rfd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_bound_signatures():
    assert declared_symbols() == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_status_strings_without_device():
    L = _lib.lib()
    assert L.rfd_status_string(0) == b"ok"
    assert L.rfd_status_string(1) == b"invalid argument"


def test_default_opts_are_the_reference_stop_rule():
    L = _lib.lib()
    o = _lib.Opts()
    L.rfd_default_opts(ctypes.byref(o))
    assert o.stop_mode == _lib.STOP_ABS
    assert o.eps == float.fromhex("0x1p-23")  # FLT_EPSILON, compute_distance_wrapper.cpp:6
    assert o.vefd_gate == 1
