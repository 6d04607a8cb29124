"""Runs the same scene through the oracle and through the gfx950 library.

A scene is a list of operations applied in order to a W x H context:
  ("clear_color", Color) / ("clear_depth", float)
  ("set", {field: value})          Context fields (ReadDepth, Cull, ...)
  ("shader", shader)               a fauxgl_amd.api built-in shader
  ("tris", Prims) / ("lines", Prims) / ("mesh", Mesh)
Both backends return (color (H,W,4) uint8, depth (H,W) float64, [infos]).
"""
import numpy as np

import oracle as O
from fauxgl_amd import api as A

_TEX_KIND = {"NRGBA": O.TEX_NRGBA8, "RGBA": O.TEX_RGBA8, "RGBA16": O.TEX_RGBA16}
_FIELDS = {"ReadDepth": "read_depth", "WriteDepth": "write_depth", "WriteColor": "write_color",
           "AlphaBlend": "alpha_blend", "Wireframe": "wireframe", "FrontFace": "front_face", "Cull": "cull",
           "LineWidth": "line_width", "DepthBias": "depth_bias"}


def oracle_shader(sh):
    if isinstance(sh, A.SolidColorShader):
        return O.make_shader(O.SHADER_SOLID, sh.Matrix, color=sh.Color)
    if isinstance(sh, A.TextureShader):
        t = O.Texture(sh.Texture.Image, _TEX_KIND[sh.Texture.kind])
        return O.make_shader(O.SHADER_TEXTURE, sh.Matrix, texture=t)
    if isinstance(sh, A.PhongShader):
        t = O.Texture(sh.Texture.Image, _TEX_KIND[sh.Texture.kind]) if sh.Texture is not None else None
        return O.make_shader(O.SHADER_PHONG, sh.Matrix, light_dir=sh.LightDirection, camera_pos=sh.CameraPosition,
                             object_color=sh.ObjectColor, ambient=sh.AmbientColor, diffuse=sh.DiffuseColor,
                             specular=sh.SpecularColor, specular_power=sh.SpecularPower, texture=t)
    raise TypeError(sh)


def _mesh_of(op, val):
    if op == "tris":
        return O.Mesh(tris=val.arrays())
    if op == "lines":
        return O.Mesh(lines=val.arrays())
    return O.Mesh(tris=val.Triangles.arrays() if val.Triangles is not None else None,
                  lines=val.Lines.arrays() if val.Lines is not None else None)


def run_oracle(W, H, ops, nthreads=1):
    ctx = O.Context(W, H)
    state = dict(read_depth=True, write_depth=True, write_color=True, alpha_blend=True, wireframe=False,
                 front_face=O.FACE_CCW, cull=O.CULL_BACK, line_width=2.0, depth_bias=0.0)
    ctx.set_state(O.make_state(**state))
    infos = []
    keep = []
    for op, val in ops:
        if op == "clear_color":
            ctx.clear_color(tuple(val))
        elif op == "clear_depth":
            ctx.clear_depth(val)
        elif op == "set":
            for k, v in val.items():
                state[_FIELDS[k]] = v
            ctx.set_state(O.make_state(**state))
        elif op == "shader":
            s = oracle_shader(val)
            keep.append(s)
            ctx.set_shader(s)
        elif op in ("tris", "lines", "mesh"):
            m = _mesh_of(op, val)
            keep.append(m)
            if op == "tris":
                infos.append(ctx.draw_triangles(m, nthreads))
            elif op == "lines":
                infos.append(ctx.draw_lines(m, nthreads))
            else:
                infos.append(ctx.draw_mesh(m, nthreads))
        else:
            raise ValueError(op)
    return ctx.color(), ctx.depth(), infos, ctx


def run_gpu(W, H, ops, resident=False):
    ctx = A.NewContext(W, H)
    infos = []
    for op, val in ops:
        if op == "clear_color":
            ctx.ClearColorBufferWith(val)
        elif op == "clear_depth":
            ctx.ClearDepthBufferWith(val)
        elif op == "set":
            for k, v in val.items():
                setattr(ctx, k, v)
        elif op == "shader":
            ctx.Shader = val
        elif op == "tris":
            infos.append(tuple(ctx.DrawTriangles(val)))
        elif op == "lines":
            infos.append(tuple(ctx.DrawLines(val)))
        elif op == "mesh":
            m = ctx.Upload(val) if resident else val
            infos.append(tuple(ctx.DrawMesh(m)))
        else:
            raise ValueError(op)
    return ctx.Image(), ctx.DepthBuffer, infos, ctx


def compare(ref, got, color_tol=1):
    """Parity: depth bit-exact, coverage (changed pixels) exact, color within
    color_tol LSB, RasterizeInfo equal.  Returns a dict of mismatch counts."""
    rc, rd, ri = ref[:3]
    gc, gd, gi = got[:3]
    out = {}
    out["info_ref"], out["info_gpu"] = ri, gi
    out["info_equal"] = list(ri) == list(gi)
    out["depth_bits_diff"] = int(np.count_nonzero(rd.view(np.uint64) != gd.view(np.uint64)))
    cdiff = np.abs(rc.astype(np.int16) - gc.astype(np.int16))
    out["color_max_diff"] = int(cdiff.max()) if cdiff.size else 0
    out["color_pixels_diff"] = int(np.count_nonzero(cdiff.max(axis=2)))
    out["ok"] = out["info_equal"] and out["depth_bits_diff"] == 0 and out["color_max_diff"] <= color_tol
    return out
