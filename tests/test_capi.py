"""The HIP library loads on a CPU-only host and exports every symbol that
include/phcmd.h declares (no compute calls without a GPU)."""
import os
import re

from phcpack_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            syms |= set(re.findall(r"\b(phc_\w+)\s*\(", text))
    return syms


def test_library_exports_declared_symbols():
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) > 10
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert declared_symbols() <= set(_lib.SIGNATURES)


def test_version_and_error_string():
    lib = _lib.load()
    assert lib.phc_version() >= 100
    assert isinstance(lib.phc_last_error(), bytes)
