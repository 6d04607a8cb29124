"""GPU check of SavePNG's encoding (fgl_encode_png): the file decodes (PIL,
an independent decoder) to the context's pixels, the IDAT stream inflates
(zlib) to exactly the filtered rows the oracle's restatement of image/png
produces, and the header and CRCs are well formed.  Each case prints its
name before touching the GPU."""
import io
import struct
import sys
import zlib

import numpy as np
import pytest
from PIL import Image

import oracle as O
from fauxgl_amd import api as A
from fauxgl_amd import meshes as M

pytestmark = pytest.mark.gpu


def _say(name):
    print(f"[case] {name}", flush=True)
    sys.stdout.flush()


def chunks(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    p, out = 8, []
    while p < len(data):
        n = struct.unpack(">I", data[p:p + 4])[0]
        kind = data[p + 4:p + 8]
        body = data[p + 8:p + 8 + n]
        crc = struct.unpack(">I", data[p + 8 + n:p + 12 + n])[0]
        assert crc == zlib.crc32(kind + body), kind
        out.append((kind, body))
        p += 12 + n
    return out


def check(ctx):
    img = ctx.Image()
    data = ctx.EncodePNG()
    h, w = img.shape[:2]
    ch = chunks(data)
    assert [k for k, _ in ch] == [b"IHDR", b"IDAT", b"IEND"]
    bpp, want = O.png_filter(img)
    assert ch[0][1] == struct.pack(">IIBBBBB", w, h, 8, 6 if bpp == 4 else 2, 0, 0, 0)
    stream = np.frombuffer(zlib.decompress(ch[1][1]), dtype=np.uint8)
    assert np.array_equal(stream, want)
    dec = np.asarray(Image.open(io.BytesIO(data)))
    assert dec.shape == (h, w, bpp) and np.array_equal(dec, img[..., :bpp])
    return len(data), len(zlib.compress(want.tobytes(), 6))


@pytest.mark.parametrize("h,w,seed", [(1, 1, 1), (2, 3, 2), (37, 53, 3), (64, 640, 4)])
def test_random_images(h, w, seed):
    _say(f"png random {w}x{h}")
    rng = np.random.Generator(np.random.PCG64(seed))
    img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    ctx = A.NewContext(w, h)
    ctx.SetColorBuffer(img)
    check(ctx)
    img[..., 3] = 255
    ctx.SetColorBuffer(img)
    check(ctx)


def test_pending_clear():
    _say("png pending clear")
    ctx = A.NewContext(33, 17)
    ctx.ClearColorBufferWith(A.HexColor("#FFF8E3"))
    check(ctx)


def test_rendered_frame():
    """The headline frame at 1080p (1x): long flat runs, Phong gradients."""
    _say("png noise sphere 1080p")
    ctx = A.NewContext(1920, 1080)
    ctx.ClearColorBufferWith(M.BACKGROUND)
    ctx.Shader = M.phong_shader(1920, 1080)
    ctx.DrawMesh(M.noise_sphere_mesh())
    size, zsize = check(ctx)
    print(f"1080p frame: PNG {size} bytes (zlib level 6 on the same rows: {zsize})", flush=True)
