"""The C-ABI library loads and exports every entry point include/ssdf_c.h declares (CPU only, no
compute calls).  Also: the oracle is absent from the product library's dependencies."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "ssdf_c.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ssdf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_reference_symbols():
    names = declared_functions()
    # the two symbols the reference exports (serial/ssdf.cpp:36, 52)
    assert "ssdf_edf_f" in names and "ssdf_edf_d" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    import ssdf_amd
    path = ssdf_amd._LIB_PATH
    assert path.exists(), "libssdf_amd.so is not built"
    lib = ctypes.CDLL(str(path))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(path)], capture_output=True, text=True).stdout
    for n in declared_functions():
        assert re.search(rf"\bT {n}$", out, re.M), n


def test_library_does_not_link_the_oracle():
    import ssdf_amd
    out = subprocess.run(["ldd", str(ssdf_amd._LIB_PATH)], capture_output=True, text=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", str(ssdf_amd._LIB_PATH)], capture_output=True, text=True).stdout
    assert "orc_" not in syms


def test_python_mirror_signatures_cover_header():
    import ssdf_amd
    missing = [n for n in declared_functions() if n not in ssdf_amd._SIGS]
    assert not missing, missing


def test_cpp_dropin_programs_build():
    """The C++ header compiles as a reference caller uses it, and links against the library."""
    from ssdf_amd import build as B
    exes = B.build_cpp_tests()
    assert exes and all(e.exists() for e in exes)
