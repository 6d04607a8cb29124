"""Golden scenes: each is built from the arrays in golden.npz (or the
build's seeded generators) and rendered by the oracle (expected.json was
made that way) or by the GPU library; tests compare hashes of the buffers.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

from fauxgl_amd import api as A  # noqa: E402
from fauxgl_amd import meshes as M  # noqa: E402
from helpers import run_gpu, run_oracle  # noqa: E402,F401

SCENES = ["hello_400x200", "cube_wireframe_160x120", "square_texture_blend_64", "sphere_cube_512"]


def load_inputs():
    with np.load(os.path.join(HERE, "golden.npz")) as z:
        return {k: z[k] for k in z.files}


def build(name, inputs):
    """(W, H, ops) of a scene (ops as in tests/helpers.py)."""
    if name == "hello_400x200":
        # examples/hello.go at 400 x 200: hello.stl, BiUnitCube,
        # SmoothNormalsThreshold(30 deg), Phong, black clear
        W, H = 400, 200
        eye, center, up = A.V(-1, -2, 2), A.V(-0.07, 0, 0), A.V(0, 0, 1)
        m = A.LookAt(eye, center, up).Perspective(20, W / H, 1, 50)
        sh = A.NewPhongShader(m, A.V(-2, 0, 1).Normalize(), eye)
        sh.ObjectColor = A.Color(0.5, 1, 0.65, 1)
        mesh = A.Mesh(A.Prims(inputs["hello_pos"], inputs["hello_nrm"]))
        return W, H, [("clear_color", A.Black), ("shader", sh), ("mesh", mesh)]
    if name == "cube_wireframe_160x120":
        # examples/cube.stl filled, then its wireframe on top with a depth
        # bias, then its edges again as lines of width 3
        W, H = 160, 120
        m = A.LookAt(A.V(3, 2.5, 4), A.V(0.5, 0.5, 0.5), A.V(0, 1, 0)).Perspective(35, W / H, 1, 20)
        pos = inputs["cube_pos"]
        tris = A.Prims(pos, inputs["cube_nrm"])
        edges = np.stack([pos[:, [0, 1]], pos[:, [1, 2]], pos[:, [2, 0]]], axis=1).reshape(-1, 2, 3)
        return W, H, [
            ("clear_color", A.HexColor("#202830")),
            ("shader", A.NewSolidColorShader(m, A.Color(0.3, 0.6, 0.9, 1))),
            ("tris", tris),
            ("set", {"Wireframe": True, "DepthBias": -1e-3, "LineWidth": 2.0}),
            ("shader", A.NewSolidColorShader(m, A.Color(1, 1, 1, 1))),
            ("tris", tris),
            ("set", {"Wireframe": False, "DepthBias": -2e-3, "LineWidth": 3.0}),
            ("shader", A.NewSolidColorShader(m, A.Color(1, 0.4, 0.1, 0.6))),
            ("lines", A.Prims(edges, nv=2)),
        ]
    if name == "square_texture_blend_64":
        # examples/square.obj with a seeded 8x8 NRGBA texture (mixed alpha),
        # drawn twice with alpha blending, the second copy shifted
        W, H = 64, 64
        m = A.Translate(A.V(-0.5, -0.5, 0)).Scale(A.V(1.6, 1.6, 1))
        tex = A.NewImageTexture(inputs["square_tex"], "NRGBA")
        pos = inputs["square_pos"]
        uv = inputs["square_uv"]
        shifted = pos + np.array([0.15, 0.1, -0.2])
        return W, H, [
            ("clear_color", A.HexColor("#FFF8E3")),
            ("set", {"Cull": A.CullNone}),
            ("shader", A.NewTextureShader(m, tex)),
            ("tris", A.Prims(pos, None, uv)),
            ("tris", A.Prims(shifted, None, uv[:, ::-1])),
        ]
    if name == "sphere_cube_512":
        # NewSphere(5) + scaled cube, Phong, 512 x 512
        W = H = 512
        return W, H, [("clear_color", M.BACKGROUND), ("shader", M.phong_shader(W, H)),
                      ("mesh", M.sphere_and_cube_mesh())]
    raise KeyError(name)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def summary(W, H, color, depth, infos):
    probe = [(W // 2, H // 2), (W // 3, H // 4), (W - 1, H - 1), (0, 0)]
    return {
        "size": [W, H],
        "info": [[int(x) for x in i] for i in infos],
        "color_sha256": sha(color),
        "depth_sha256": sha(depth.view(np.uint64)),
        "pixels": [[x, y] + [int(v) for v in color[y, x]] for x, y in probe],
    }
