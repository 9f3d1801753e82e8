"""The C-ABI library loads and exports every entry point include/pfdtd.h declares (no compute
calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

from parallelfdtd_amd import native

HEADER = Path(__file__).resolve().parent.parent / "include" / "pfdtd.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(pfdtd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "pfdtd_run" in names and "pfdtd_ctx_create" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = native.lib()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared()) <= set(native.SIGNATURES)


def test_error_path_without_device():
    lib = native.lib()
    rc = lib.pfdtd_ctx_destroy(None)
    assert rc == native.OK
    rc = lib.pfdtd_set_kernel_timing(None, 1)
    assert rc == native.E_ARG
    assert b"null" in lib.pfdtd_last_error()
