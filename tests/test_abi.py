"""The C ABI: the built library exports every entry point include/fgl.h
declares, and the Python bindings lay out the header's structs the same way
a C compiler does.  CPU only (nothing here calls into the device)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from fauxgl_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fgl.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(fgl_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_draw_path():
    names = declared_functions()
    for must in ("fgl_ctx_create", "fgl_draw_triangles", "fgl_draw_lines", "fgl_draw_mesh", "fgl_read_color",
                 "fgl_read_depth", "fgl_clear_color", "fgl_clear_depth", "fgl_resolve"):
        assert must in names


def test_bindings_cover_the_header():
    assert sorted(N.EXPORTED) == declared_functions()


def test_library_exports_every_declared_symbol():
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(N.LIB_PATH)
    missing = [name for name in declared_functions() if not hasattr(lib, name)]
    assert not missing, missing


def test_library_exports_nothing_else_in_the_fgl_namespace():
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("library not built")
    out = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("nm unavailable")
    exported = sorted({ln.split()[-1] for ln in out.stdout.splitlines() if re.search(r"\bT fgl_", ln)})
    assert exported == declared_functions()


STRUCTS = {
    "fgl_state": (N.FglState, ["read_depth", "write_depth", "write_color", "alpha_blend", "wireframe",
                               "front_face", "cull", "line_width", "depth_bias"]),
    "fgl_shader": (N.FglShader, ["kind", "matrix", "color", "light_dir", "camera_pos", "object_color", "ambient",
                                 "diffuse", "specular", "specular_power", "texture"]),
    "fgl_prims": (N.FglPrims, ["count", "position", "normal", "texture", "color"]),
    "fgl_kernel_time": (N.FglKernelTime, ["name", "launches", "total_ms"]),
    "fgl_draw_stat": (N.FglDrawStat, ["total_pixels", "updated_pixels", "clipped_triangles", "side_triangles",
                                      "pairs", "spans", "fragments", "overflow_mask", "reruns"]),
}


def test_struct_layouts_match_the_c_compiler():
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fgl.h"', "int main(void) {"]
    for cname, (_, fields) in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in fields:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    got = {}
    for ln in out.splitlines():
        parts = ln.split()
        got[tuple(parts[:-1])] = int(parts[-1])
    for cname, (cls, fields) in STRUCTS.items():
        assert ctypes.sizeof(cls) == got[(cname, "size")], cname
        for f in fields:
            assert getattr(cls, f).offset == got[(cname, f)], (cname, f)
