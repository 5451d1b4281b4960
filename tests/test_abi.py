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


def test_ctypes_structs_match_the_header_layout(tmp_path):
    """Every field of the ctypes mirrors sits at the header's offset, and the
    structs have the header's size (a field added on one side only would
    shift the others silently)."""
    import subprocess
    from dnds_amd import _capi
    pairs = {"dndsmi_mesh_view": _capi.MeshView, "dndsmi_mesh_spec": _capi.MeshSpec,
             "dndsmi_vr_setting": _capi.VRSetting, "dndsmi_vr_view": _capi.VRView,
             "dndsmi_euler_setting": _capi.EulerSetting}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dndsmi.h"', "int main(void) {"]
    for cname, py in pairs.items():
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'    printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("    return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, field, val = line.split()
        got[(cname, field)] = int(val)
    for cname, py in pairs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_oracle_ctypes_structs_match_its_header(tmp_path):
    import subprocess
    import oracle.oracle as O
    pairs = {"oracle_mesh": O.OMesh, "oracle_setup": O.OSetup, "oracle_setting": O.OSetting}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "oracle.h"', "int main(void) {"]
    for cname, py in pairs.items():
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'    printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("    return 0;\n}")
    src = tmp_path / "olayout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "olayout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "oracle"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, field, val = line.split()
        got[(cname, field)] = int(val)
    for cname, py in pairs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
