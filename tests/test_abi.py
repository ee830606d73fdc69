"""The C-ABI library loads and exports every entry point include/pfgpu.h declares (CPU only)."""
import ctypes as C
import os
import re

import pytest

from polyfem_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "pfgpu.h")).read()
    return sorted(set(re.findall(r"\b(pfgpu_[a-z_]+)\s*\(", text)))


def test_header_declares_the_binding():
    assert _declared() == _ffi.exported_symbols()


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_ffi.LIB_PATH):
        pytest.fail(f"{_ffi.LIB_PATH} not built")
    lib = C.CDLL(_ffi.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_no_oracle_in_product_path():
    """The product package never imports or loads the oracle."""
    pkg = os.path.join(ROOT, "polyfem_amd")
    banned = ("pyoracle", "liboracle", "import oracle", "from oracle", "oracle.h", "orc_")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert not any(b in txt for b in banned), f


def test_null_arguments_are_rejected_without_a_device():
    lib = _ffi.lib()
    st = lib.pfgpu_create(None, None, 0, None)
    assert st == 1
    assert b"null" in lib.pfgpu_last_error()
    assert lib.pfgpu_value(None, None, None) == 1
