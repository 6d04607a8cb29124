"""numpy PNG row filters and their inverse (RFC 2083 section 6), written
from the PNG specification for the tests."""
import numpy as np


def paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
    return np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))


def candidates(cur, prev, bpp):
    """The five filtered versions (None, Sub, Up, Average, Paeth) of a row."""
    cur = cur.astype(np.int64)
    prev = prev.astype(np.int64)
    a = np.concatenate([np.zeros(bpp, np.int64), cur[:-bpp]])
    c = np.concatenate([np.zeros(bpp, np.int64), prev[:-bpp]])
    b = prev
    return [cur % 256, (cur - a) % 256, (cur - b) % 256, (cur - (a + b) // 2) % 256, (cur - paeth(a, b, c)) % 256]


def unfilter(stream, w, h, bpp):
    """Filtered stream -> (h, w * bpp) raw rows."""
    n = w * bpp
    rows = np.zeros((h, n), np.int64)
    prev = np.zeros(n, np.int64)
    for y in range(h):
        f = int(stream[y * (n + 1)])
        d = stream[y * (n + 1) + 1:(y + 1) * (n + 1)].astype(np.int64)
        cur = np.zeros(n, np.int64)
        for i in range(n):
            a = cur[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) // 2, int(paeth(np.int64(a), np.int64(b), np.int64(c)))][f]
            cur[i] = (d[i] + pred) % 256
        rows[y] = cur
        prev = cur
    return rows
