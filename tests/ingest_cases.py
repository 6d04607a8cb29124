"""Seeded inputs for the mesh-ingest tests (LoadSTL / LoadOBJ): binary and
ASCII STL and OBJ files built in memory, plus the reference's example
assets kept as fixtures in tests/golden (cube.stl binary, hello.stl ASCII,
square.obj)."""
import os
import struct

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture(name):
    with open(os.path.join(GOLDEN, name), "rb") as f:
        return f.read()


def binary_stl(tris32, header=b"seeded", attr=0):
    """stl.go's binary layout: 80-byte header, uint32 count, 50-byte records
    (normal, three vertices as float32, uint16 attribute)."""
    tris32 = np.asarray(tris32, dtype=np.float32).reshape(-1, 3, 3)
    out = bytearray(header[:80].ljust(80, b"\0"))
    out += struct.pack("<I", len(tris32))
    for t in tris32:
        out += struct.pack("<3f", 0.0, 0.0, 0.0)
        out += t.astype("<f4").tobytes()
        out += struct.pack("<H", attr)
    return bytes(out)


def random_binary_stl(seed, n=500):
    rng = np.random.Generator(np.random.PCG64(seed))
    t = rng.normal(size=(n, 3, 3)).astype(np.float32) * 3
    t[0, 1] = t[0, 0]  # degenerate: NaN normal
    t[1] = 0.0
    t[2, 0, 0] = -0.0
    return t, binary_stl(t)


def ascii_stl(verts, tokens=None):
    """ASCII STL; tokens overrides the text of chosen coordinates."""
    lines = ["solid seeded"]
    for i in range(0, len(verts), 3):
        lines += ["  facet normal 0 0 0", "    outer loop"]
        for v in verts[i:i + 3]:
            lines.append("      vertex " + " ".join(v))
        lines += ["    endloop", "  endfacet"]
    lines.append("endsolid seeded")
    return ("\r\n".join(lines) + "\r\n").encode()


ODD_TOKENS = [
    ("1e5", 1e5), (".5", 0.5), ("5.", 5.0), ("-0", -0.0), ("+2.5", 2.5), ("inf", np.inf), ("-Infinity", -np.inf),
    ("1e", 0.0), ("abc", 0.0), ("1_0", 0.0), ("0x1p-2", 0.25), ("0x1.8P1", 3.0), ("0x1", 0.0), ("1e400", np.inf),
    ("4.9e-324", 5e-324), ("0.1", 0.1), ("1.5.2", 0.0), ("--1", 0.0), ("", None),
]


def odd_ascii_stl():
    """One vertex per token of ODD_TOKENS (x), with plain y, z; returns the
    file and the expected x values."""
    toks = [(t, v) for t, v in ODD_TOKENS if t]
    while len(toks) % 3:
        toks.append(("7", 7.0))
    verts = [[t, "1.25", repr(0.1 * i)] for i, (t, _) in enumerate(toks)]
    return ascii_stl(verts), np.array([v for _, v in toks]), np.array([0.1 * i for i in range(len(toks))])


def random_obj(seed, nv=60, nf=80):
    """Polygons of 3..6 corners over nv vertices, with uv and normals (some
    zero, so FixNormals applies), mixed index forms: 'a', 'a/b', 'a//c',
    'a/b/c', negative (relative) indices, hexadecimal and octal indices."""
    rng = np.random.Generator(np.random.PCG64(seed))
    v = rng.normal(size=(nv, 3))
    vt = rng.uniform(size=(nv, 2))
    vn = rng.normal(size=(nv, 3))
    vn[::7] = 0.0
    lines = ["# seeded", "o thing"]
    lines += [f"v {float(a)!r} {float(b)!r} {float(c)!r}" for a, b, c in v]
    lines += [f"vt {float(a)!r} {float(b)!r}" for a, b in vt]
    lines += [f"vn {float(a)!r} {float(b)!r} {float(c)!r}" for a, b, c in vn]
    for k in range(nf):
        m = int(rng.integers(3, 7))
        corners = rng.integers(1, nv + 1, size=m)
        parts = []
        for c in corners:
            form = int(rng.integers(0, 6))
            if form == 0:
                parts.append(f"{c}")
            elif form == 1:
                parts.append(f"{c}/{c}")
            elif form == 2:
                parts.append(f"{c}//{c}")
            elif form == 3:
                parts.append(f"{c}/{c}/{c}")
            elif form == 4:
                parts.append(f"{c - nv - 1}/{c - nv - 1}/{c - nv - 1}")  # relative
            else:
                parts.append(f"{hex(c)}/0o{c:o}/{c}")
        lines.append("f " + " ".join(parts))
    lines.append("s off")
    return ("\n".join(lines) + "\n").encode()
