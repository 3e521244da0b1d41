"""The C-ABI libraries load and export every function their headers declare."""
import ctypes
import os
import re

import pytest

from fastercap_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:fcg|fch)_[a-z_0-9]+)\s*\(", text)))


@pytest.mark.parametrize("header,loader,listed", [
    ("fastercap_amd.h", _ffi.device_lib, _ffi.DEVICE_SYMBOLS),
    ("fastercap_host.h", _ffi.host_lib, _ffi.HOST_SYMBOLS),
])
def test_exports(header, loader, listed):
    names = declared(header)
    assert names, header
    assert sorted(names) == sorted(listed)
    lib = loader()
    for n in names:
        assert hasattr(lib, n), n


def test_node_struct_size():
    assert ctypes.sizeof(_ffi.NodeGeom) == 96


def _defines(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return {m.group(1): int(m.group(2)) for m in re.finditer(r"#define\s+(FCG_\w+)\s+(\d+)u?\b", text)}


# Values of the reference's own macros, transcribed with their file:line:
#   AUTOREFINE_PRECOND_NONE/JACOBI/BLOCK/SUPER/HIER  Solver/SolverGlobal.h:60-68
#   FC_NORMAL_END .. FC_USER_BREAK                    Solver/SolverGlobal.h:39-49
#   AUTOPANEL_IS_LEAF .. AUTOPANEL_IS_SUPER_NODE      Solver/AutoElement.h:44-48
#   AUTOREFINE_COLLOCATION / AUTOREFINE_GALERKIN      Solver/SolverGlobal.h:72-73
REFERENCE_VALUES = {
    "FCG_SCHEME_COLLOCATION": 1, "FCG_SCHEME_GALERKIN": 2,
    "FCG_PRECOND_NONE": 0, "FCG_PRECOND_JACOBI": 1, "FCG_PRECOND_BLOCK": 2, "FCG_PRECOND_SUPER": 4,
    "FCG_PRECOND_HIER": 8,
    "FCG_OK": 0, "FCG_GENERIC_ERROR": 1, "FCG_COMMAND_LINE_ERROR": 64, "FCG_CANNOT_OPEN_FILE": 66,
    "FCG_OUT_OF_MEMORY": 71, "FCG_FILE_ERROR": 74, "FCG_EXCEPTION_ERROR": 98, "FCG_USER_BREAK": 125,
    "FCG_NODE_LEAF": 1, "FCG_NODE_DIEL": 2, "FCG_NODE_OUTPERM_NORMAL_DIR": 4, "FCG_NODE_OUTPERM_ELEMENT_LEVEL": 8,
    "FCG_NODE_SUPER": 128,
}


def test_boundary_values_are_the_references():
    """A caller of the reference passes m_ucPrecondType, FC_* codes and m_ucType
    flags unchanged: the header's values must be the reference's."""
    d = _defines("fastercap_amd.h")
    for name, val in REFERENCE_VALUES.items():
        assert d.get(name) == val, (name, d.get(name), val)
        if hasattr(_ffi, name):
            assert getattr(_ffi, name) == val, name
    # the precond bit set of the reference combines flags: each flag is one bit
    for name in ("FCG_PRECOND_JACOBI", "FCG_PRECOND_BLOCK", "FCG_PRECOND_SUPER", "FCG_PRECOND_HIER"):
        v = d[name]
        assert v & (v - 1) == 0


def test_cli_precond_values():
    """The Python command line stores the reference's bit values."""
    from fastercap_amd import cli

    assert cli.parse_args(["-ps256", "x.txt"]).precond == _ffi.FCG_PRECOND_SUPER
    assert cli.parse_args(["-pb16", "x.txt"]).precond == _ffi.FCG_PRECOND_BLOCK
    assert cli.parse_args(["-ps256", "-pj", "x.txt"]).precond == _ffi.FCG_PRECOND_SUPER
    assert cli.auto_preconditioner(10_000_000, 4)[0] == _ffi.FCG_PRECOND_SUPER
    assert cli.auto_preconditioner(1000, 1)[0] == _ffi.FCG_PRECOND_JACOBI


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes mirror of a header struct has the C compiler's size and field
    offsets (a C program built from include/fastercap_amd.h prints them)."""
    structs = {
        "fcg_solve_opts": _ffi.SolveOpts,
        "fcg_solve_stats": _ffi.SolveStats,
        "fcg_job_mem": _ffi.JobMem,
        "fcg_mvp_timing": _ffi.MvpTiming,
        "fcg_profile": _ffi.Profile,
        "fcg_node_geom": _ffi.NodeGeom,
        "fcg_job_desc": _ffi.JobDesc,
    }
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "fastercap_amd.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    import subprocess

    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for ln in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        c, f, v = ln.split()
        got[(c, f)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
