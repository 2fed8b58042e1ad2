"""The C-ABI library loads and exports every entry point include/mdmc.h declares (no GPU needed)."""
import ctypes
import os
import re

import manifolddmc_amd as mdmc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "mdmc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mdmc_[a-z_]+)\s*\(", src)))


def test_header_declares_the_python_exports():
    assert declared_functions() == sorted(mdmc.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(mdmc.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_options_struct_layout_matches_header():
    assert ctypes.sizeof(mdmc._Options) == 16
    assert ctypes.sizeof(mdmc._Grid) == 3 * 4 * 3 + 4 + 8 + 4 + 4  # dims, origin, delta, pad, ptr, flag, tail pad


def test_struct_mirrors_match_the_compiled_header(tmp_path):
    """Every field of the ctypes mirrors sits at the offset the C compiler gives it in
    include/mdmc.h, and the structs have the same size (a field added on one side only
    shifts everything after it)."""
    import shutil
    import subprocess

    import pytest
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    structs = [("mdmc_options", mdmc._Options), ("mdmc_grid", mdmc._Grid), ("mdmc_stats", mdmc.Stats)]
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "mdmc.h"', "int main(void) {"]
    for cname, py in structs:
        lines.append('    printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in py._fields_:
            lines.append('    printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        a, b, v = line.split()
        got[(a, b)] = int(v)
    for cname, py in structs:
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])


def test_reference_member_names_map_to_flags():
    # DualMarchingCubes.cu:110-133 defaults; the reference names alias the snake_case options
    d = object.__new__(mdmc.DualMarchingCubes)  # no context: flags() is host logic only
    d.enable_surface_to_vertex_iteration = True
    d.simplify_p3x3y_color = d.simplify_p3x3y_old = d.simplify_p3333 = False
    d.split_non_manifold_halfedges = d.detect_non_manifold_vertices = True
    d.splitting_step = mdmc.SplittingStep.ALL
    assert d.flags() == mdmc.DEFAULT_FLAGS
    d.m_simplifyP3x3yColor = True
    d.m_simplifyP3333 = True
    d.m_enableSurfaceToVertexIteration = False
    d.m_programmaticSplittingStep = mdmc.SplittingStep.MERGE_VERTICES
    assert d.simplify_p3x3y_color and d.simplify_p3333 and not d.enable_surface_to_vertex_iteration
    assert d.splitting_step == mdmc.SplittingStep.MERGE_VERTICES
    assert d.flags(add_zero_boundary=False) == (mdmc.P3X3Y_COLOR | mdmc.P3333 | mdmc.SPLIT_NM | mdmc.DETECT_NM_VERTS)
