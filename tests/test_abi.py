"""The C-ABI library loads and exports every entry point the header declares
(no device calls: runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "dndsmi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dndsmi_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "dndsmi_reconstruction_jacobi_step" in syms
    assert "dndsmi_evaluate_rhs" in syms
    assert len(syms) >= 20


def test_library_exports_every_declared_symbol():
    from dnds_amd import _capi
    lib = ctypes.CDLL(_capi.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_signatures_cover_header():
    from dnds_amd import _capi
    declared = set(declared_symbols())
    bound = {n for n, _, _ in _capi.SIGNATURES}
    assert declared == bound, (declared - bound, bound - declared)


def test_oracle_is_separate_library():
    import oracle
    from dnds_amd import _capi
    lib = ctypes.CDLL(_capi.LIB_PATH)
    assert not hasattr(lib, "oracle_jacobi_step")
    olib = ctypes.CDLL(oracle.LIB)
    assert hasattr(olib, "oracle_jacobi_step")
