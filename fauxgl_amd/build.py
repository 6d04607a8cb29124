"""Builds libfauxgl_amd.so (the gfx950 draw path and its C ABI) in-tree.

Run as ``python -m fauxgl_amd.build`` or through ``__graft_entry__.build()``.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfauxgl_amd.so")

SOURCES = ["fgl_vs.hip", "fgl_bin.hip", "fgl_spans.hip", "fgl_tile.hip", "fgl_misc.hip", "fgl_nn.hip", "fgl_mesh.hip", "fgl_ingest.hip", "fgl_png.hip", "fgl_host.cpp"]

# -ffp-contract=off: Go evaluates a*b+c as two roundings; the FP64 coverage
# and depth arithmetic must match it bit for bit.  Packed FP32 ops are
# disabled for all device code.
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the gfx950 library cannot be built")


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libfauxgl_amd.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
