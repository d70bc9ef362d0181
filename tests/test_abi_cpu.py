"""CPU checks of the C ABI: the library loads and exports every entry point
include/sparsebackends_amd.h declares (no device calls)."""
import os
import re

import sparsebackends_amd as sb
from sparsebackends_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "sparsebackends_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sb_[a-z0-9_]+)\s*\(", txt, re.M)))


def test_header_matches_binding_table():
    assert declared() == sorted(_abi.SIGNATURES)


def test_library_exports_every_symbol():
    lib = sb.load()
    for name in declared():
        assert hasattr(lib, name), name


def test_version_and_error_string():
    lib = sb.load()
    assert lib.sb_version() >= 10000
    assert isinstance(lib.sb_last_error(), bytes)
