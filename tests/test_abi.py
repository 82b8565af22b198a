"""CPU checks of the C-ABI boundary (no GPU compute calls)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import qsolver as Q
from kat_cases import splitmix64_uniform

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qsolver.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qs_[A-Za-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = Q.load()
    names = declared_functions()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers every declared entry point
    assert sorted(Q._SIGS) == names


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "qsolver.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(Mesh), offsetof(Mesh, Lx),"
        " offsetof(Mesh, specsize), offsetof(Mesh, spec), offsetof(Mesh, priv), sizeof(qs_params),"
        " offsetof(qs_params, advect_w)); return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    want = [ctypes.sizeof(Q.MeshStruct), Q.MeshStruct.Lx.offset, Q.MeshStruct.specsize.offset,
            Q.MeshStruct.spec.offset, Q.MeshStruct.priv.offset, ctypes.sizeof(Q.QsParams),
            Q.QsParams.advect_w.offset]
    assert got == want


def test_cpp_wrapper_header_builds_and_links(tmp_path):
    """include/qsolver.hpp (reference-named C++ wrappers) compiles against the
    C header and links against libqsolver.so (nothing is run)."""
    src = tmp_path / "port.cpp"
    src.write_text(
        '#include "qsolver.hpp"\n'
        "int main(){ Mesh* m = qsolver::make_mesh(64, 64, 6.283185307179586, 6.283185307179586);\n"
        " Field* f = qsolver::make_field(m); qsolver::FwdTrans(f->spec, f->phys, m);\n"
        " qsolver::IFRK2::integrate_func0D(f->spec, f->spec, f->spec, m->k_squared, m->k_squared, 1e-3, m);\n"
        " qs_params p{}; p.lambda = p.Re = p.Er = p.alpha = p.D_Q = p.a_Q = p.c_Q = 1; p.dt = 1e-3;\n"
        " qsolver::Solver s(m, p); s.step_ifrk2(1); qs_field_destroy(f); qs_mesh_destroy(m); return 0; }\n")
    lib = os.path.dirname(Q.load()._name)
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                           str(src), "-o", str(tmp_path / "port"), "-L", lib, "-lqsolver",
                           "-Wl,-rpath," + lib, "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib"])


def test_synth_uniform_matches_splitmix64():
    got = Q.synth_uniform(12345, 0.0, 1.0, 1000)
    assert np.array_equal(got, splitmix64_uniform(12345, 1000))
    got = Q.synth_uniform(1, -0.01, 0.01, 10)
    assert np.array_equal(got, -0.01 + 0.02 * splitmix64_uniform(1, 10))


def test_errors_are_status_codes_not_aborts():
    L = Q.load()
    m = ctypes.POINTER(Q.MeshStruct)()
    assert L.qs_mesh_create(100, 64, 1.0, 1.0, 16, ctypes.byref(m)) == 2   # not a power of two
    assert b"Nx=100" in L.qs_last_error()
    assert L.qs_mesh_create(64, 64, -1.0, 1.0, 16, ctypes.byref(m)) == 1
    assert L.qs_mesh_create(64, 64, 1.0, 1.0, 16, None) == 1
    assert L.qs_FwdTrans(None, None, None) == 1
    assert L.qs_step_ifrk2(None, 1) == 1
    assert L.qs_status_string(4) == b"non-finite value in state"


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "nature-control_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")) or f == "Makefile":
                text = open(os.path.join(dirpath, f), errors="replace").read()
                for pat in ("oracle_py", "liboracle", "qs_oracle", "or_solver", "or_mesh"):
                    assert pat not in text, (pat, os.path.join(dirpath, f))
