"""GPU parity of the SSAA resolve (the step after the draw path) against the
oracle's restatement of the same filter: bit-exact, integer arithmetic.
Each case prints its name before touching the GPU."""
import sys

import numpy as np
import pytest

import oracle as O
from fauxgl_amd import api as A
from fauxgl_amd import meshes as M

pytestmark = pytest.mark.gpu

FILTERS = {"tent": 1, "box": 2}


def _say(name):
    print(f"[case] {name}", flush=True)
    sys.stdout.flush()


def resolve_gpu(img, dw, dh, filt, pending_clear=None):
    sh, sw = img.shape[:2]
    src = A.NewContext(sw, sh)
    dst = A.NewContext(dw, dh)
    if pending_clear is None:
        src.SetColorBuffer(img)
    else:
        src.ClearColorBufferWith(pending_clear)
    src.Resolve(dst, filt)
    return dst.Image()


def random_nrgba(h, w, seed, opaque=False):
    rng = np.random.Generator(np.random.PCG64(seed))
    img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    if opaque:
        img[..., 3] = 255
    else:
        # a mix of opaque, transparent and partial alpha
        img[..., 3] = rng.choice(np.array([0, 1, 128, 254, 255], np.uint8), (h, w))
    return img


@pytest.mark.parametrize("filt", sorted(FILTERS))
@pytest.mark.parametrize("f,dw,dh", [(1, 9, 4), (2, 17, 11), (3, 37, 23), (4, 64, 36), (4, 101, 7), (4, 1, 1),
                                     (4, 16, 15), (4, 33, 31), (5, 13, 13)])
def test_resolve_random(filt, f, dw, dh):
    _say(f"resolve {filt} x{f} {dw}x{dh}")
    img = random_nrgba(dh * f, dw * f, seed=100 + f * 7 + dw)
    got = resolve_gpu(img, dw, dh, filt)
    ref = O.resolve(img, dw, dh, FILTERS[filt])
    assert np.array_equal(got, ref), np.argwhere(got != ref)[:5]


@pytest.mark.parametrize("dw,dh", [(250, 47), (17, 46)])
def test_resolve_tent_4x_opaque_and_mixed_blocks(dw, dh):
    """4x tent where most 4-pixel groups are opaque (premultiply skipped) and a
    few scattered pixels are not, so one workgroup mixes both."""
    _say(f"resolve tent x4 {dw}x{dh} mostly opaque")
    img = random_nrgba(dh * 4, dw * 4, seed=7 + dw, opaque=True)
    rng = np.random.Generator(np.random.PCG64(dw * 31 + dh))
    ys = rng.integers(0, dh * 4, 40)
    xs = rng.integers(0, dw * 4, 40)
    img[ys, xs, 3] = rng.integers(0, 255, 40, dtype=np.uint8)
    got = resolve_gpu(img, dw, dh, "tent")
    ref = O.resolve(img, dw, dh, 1)
    assert np.array_equal(got, ref), np.argwhere(got != ref)[:5]


@pytest.mark.parametrize("filt", sorted(FILTERS))
def test_resolve_of_a_pending_clear(filt):
    _say(f"resolve {filt} of a deferred clear")
    c = A.HexColor("#FFF8E3")
    got = resolve_gpu(np.zeros((32, 48, 4), np.uint8), 12, 8, filt, pending_clear=c)
    img = np.broadcast_to(np.array(c.NRGBA(), np.uint8), (32, 48, 4)).copy()
    assert np.array_equal(got, O.resolve(img, 12, 8, FILTERS[filt]))


def test_resolve_full_16x_render():
    """The headline step: a 7680x4320 render of the noise sphere resolved to
    1920x1080, against the oracle's resolve of the same GPU render (the draw
    itself is covered by the draw-path parity tests)."""
    _say("resolve tent of the 16x noise-sphere render")
    W, H, s = 1920, 1080, 4
    big = A.NewContext(W * s, H * s)
    out = A.NewContext(W, H)
    big.Shader = M.phong_shader(W, H)
    big.ClearColorBufferWith(M.BACKGROUND)
    big.DrawMesh(M.noise_sphere_mesh(ntri=100_000))
    big.Resolve(out, "tent")
    got = out.Image()
    ref = O.resolve(big.Image(), W, H, 1)
    assert np.array_equal(got, ref)
