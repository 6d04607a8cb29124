"""fauxgl_amd: fauxgl's draw path on AMD Instinct MI355X (gfx950).

The Go API of the reference (NewContext, Context.DrawMesh / DrawTriangles /
DrawLines, the clears, the built-in shaders, Image / DepthImage) over a C ABI
(include/fgl.h) whose kernels are hand-written HIP for gfx950.
"""
from .api import (Black, Color, Context, CullBack, CullFront, CullNone, DeviceMesh, Discard, FaceCCW, FaceCW,
                  Frustum, Gray, HexColor, Identity, ImageTexture, LookAt, Matrix, MaxFloat64, Mesh, NewContext,
                  NewImageTexture, NewLineMesh, NewPhongShader, NewSolidColorShader, NewTextureShader,
                  NewTriangleMesh, Orthographic, Perspective, PhongShader, Prims, Radians, RasterizeInfo, Scale,
                  Screen, SolidColorShader, TextureShader, Translate, Transparent, V, Vector, White)

__all__ = [n for n in dir() if not n.startswith("_")]
