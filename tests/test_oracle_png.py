"""Pins the oracle's restatement of image/png's row filtering (the encoder
SavePNG, util.go:56-63, goes through): every row is one of the five PNG
filters of the raw row, the chosen one has the smallest sum of |byte| read
as int8 with ties resolved in the order Up, Paeth, None, Sub, Average, and
the stream unfilters (PNG specification) back to the image.  RGB exactly
when every alpha is 255."""
import numpy as np

import oracle as O
import png_ref as P


def abs8(v):
    v = np.asarray(v, np.int64)
    return np.where(v < 128, v, 256 - v)


def check(img):
    h, w = img.shape[:2]
    bpp, s = O.png_filter(img)
    assert bpp == (3 if (img[..., 3] == 255).all() else 4)
    raw = img[..., :bpp].reshape(h, w * bpp)
    n = w * bpp
    prev = np.zeros(n, np.int64)
    for y in range(h):
        f = int(s[y * (n + 1)])
        row = s[y * (n + 1) + 1:(y + 1) * (n + 1)]
        cands = P.candidates(raw[y], prev, bpp)
        sums = [int(abs8(cv).sum()) for cv in cands]
        order = [2, 4, 0, 1, 3]
        best = min(sums)
        assert f == next(k for k in order if sums[k] == best), (y, sums, f)
        assert np.array_equal(row, cands[f])
        prev = raw[y].astype(np.int64)
    assert np.array_equal(P.unfilter(s, w, h, bpp), raw)


def test_random_images():
    rng = np.random.Generator(np.random.PCG64(4))
    for h, w in ((1, 1), (3, 5), (9, 17)):
        img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
        check(img)
        img[..., 3] = 255
        check(img)


def test_smooth_and_flat_images():
    ys, xs = np.mgrid[0:12, 0:20]
    img = np.zeros((12, 20, 4), np.uint8)
    img[..., 0] = (xs * 9) % 256  # Sub-friendly
    img[..., 1] = (ys * 13) % 256  # Up-friendly
    img[..., 2] = 77
    img[..., 3] = 255
    check(img)
    check(np.zeros((4, 4, 4), np.uint8))  # transparent black: ties everywhere
