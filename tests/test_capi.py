"""The C-ABI library loads and exports every symbol include/clblm_hip.h declares."""
import ctypes
import os
import re

from clblm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_symbols_exported():
    hdr = open(os.path.join(ROOT, "include", "clblm_hip.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char \*|double)\s*\*?\s*(clblm_\w+)\s*\(", hdr, re.M))
    assert declared, "no declarations parsed"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(_lib.EXPORTS), declared ^ set(_lib.EXPORTS)


def test_version_without_gpu():
    lib = _lib.load()
    assert lib.clblm_version() >= 1
