"""Builds libfauxgl_amd.so (the gfx950 draw path and its C ABI) in-tree.

Run as ``python -m fauxgl_amd.build`` or through ``__graft_entry__.build()``.
"""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfauxgl_amd.so")

SOURCES = ["fgl_vs.hip", "fgl_bin.hip", "fgl_spans.hip", "fgl_tile.hip", "fgl_misc.hip", "fgl_nn.hip", "fgl_mesh.hip", "fgl_ingest.hip", "fgl_png.hip", "fgl_host.cpp"]

# The draw kernels are compiled once per tile shape (the shapes of
# FGL_TILE_SHAPES in csrc/fgl_internal.h); the other sources once.
TILE_SOURCES = ["fgl_vs.hip", "fgl_bin.hip", "fgl_spans.hip", "fgl_tile.hip"]
TILE_SHAPES = [(32, 32), (64, 32), (32, 16), (64, 64), (128, 64)]

# -ffp-contract=off: Go evaluates a*b+c as two roundings; the FP64 coverage
# and depth arithmetic must match it bit for bit.  Packed FP32 ops are
# disabled for all device code.
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def units():
    """(source, extra flags, object name) for every compilation."""
    out = []
    for src in SOURCES:
        if src in TILE_SOURCES:
            for w, h in TILE_SHAPES:
                out.append((src, [f"-DFGL_TILE_W={w}", f"-DFGL_TILE_H={h}"], f"{src}.t{w}x{h}.o"))
        else:
            out.append((src, [], f"{src}.o"))
    return out


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the gfx950 library cannot be built")


def build(verbose=False, extra=()):
    """Compiles every unit (in parallel) and links the library; `extra` adds
    flags to every compilation."""
    cc = hipcc()
    jobs = max(1, min(len(units()), os.cpu_count() or 4, 16))
    with tempfile.TemporaryDirectory(prefix="fgl_build_") as tmpdir:
        def compile_one(u):
            src, defs, obj = u
            cmd = [cc] + FLAGS + list(extra) + defs + ["-c", os.path.join(CSRC, src), "-o", os.path.join(tmpdir, obj)]
            if verbose:
                print(" ".join(cmd), flush=True)
            return cmd, subprocess.run(cmd, capture_output=True, text=True)

        with ThreadPoolExecutor(jobs) as ex:
            results = list(ex.map(compile_one, units()))
        for cmd, res in results:
            if res.returncode != 0:
                sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
                raise RuntimeError("hipcc failed building libfauxgl_amd.so")
        tmp = LIB + ".tmp"
        link = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [os.path.join(tmpdir, o) for _, _, o in units()]
        if verbose:
            print(" ".join(link), flush=True)
        res = subprocess.run(link, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("linking libfauxgl_amd.so failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
