"""The C ABI boundary (CPU only): the shared library loads, exports every
entry point include/dmrg_hip.h declares, and the Python binding declares
exactly that set.  No compute call is made here (there is no GPU)."""

import ctypes
import os
import re

from dmrgpy_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "dmrg_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dmrg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    lib = _lib.load()
    names = header_functions()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name


def test_binding_matches_header():
    assert sorted(_lib.SIGNATURES) == header_functions()


def test_struct_layouts():
    # 10 ints then pointers: the C layout puts the first pointer at offset 40
    assert _lib.HeffDesc.L.offset == 40
    assert ctypes.sizeof(_lib.HeffDesc) == 40 + 4 * 8 + 8
    assert _lib.EnvDesc.ket.offset == 40
    assert ctypes.sizeof(_lib.EnvDesc) == 40 + 3 * 8 + 8 + 2 * 8


def test_version_and_error_are_callable_without_gpu():
    lib = _lib.load()
    assert lib.dmrg_version() >= 100
    assert lib.dmrg_last_error() is not None
