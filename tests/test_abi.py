"""The C ABI library loads and exports every symbol declared in
include/scico_amd.h, and the ctypes signature table matches the header
(CPU only: no compute calls)."""

import ctypes
import os
import re

from scico_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "scico_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(scxr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    syms = header_symbols()
    assert len(syms) >= 10
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_table_matches_header():
    bound = sorted(name for name, _, _ in _lib.SIGNATURES)
    assert bound == header_symbols()


def test_version_and_error_without_gpu():
    lib = _lib.load()
    assert lib.scxr_version() == 1
    assert isinstance(lib.scxr_last_error(), bytes)
