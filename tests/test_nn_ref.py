"""Pins the CPU reference of the learned anti-aliasing network (tests/nn_ref.py)
that the GPU tests compare against: its convolution layout against explicit
loops over pixels and taps, its gradient against central differences in
float64, its BF16 rounding against the definition, its features against a
hand-computed pixel."""
import numpy as np
import torch

from fauxgl_amd import nnaa
import nn_ref as R


def rand_params(seed, scale=0.2):
    _, n = nnaa.layer_offsets()
    return np.random.Generator(np.random.PCG64(seed)).normal(0, scale, n).astype(np.float32)


def loop_forward(params, feat):
    """out[y][x][co] = b[co] + sum over taps and ci of w[tap][co][ci] *
    in[y + tap // 3 - 1][x + tap % 3 - 1][ci] (zero outside), ReLU between."""
    x = feat.astype(np.float64)
    H, W, _ = x.shape
    layers = nnaa.split_params(params)
    for li, (w, b) in enumerate(layers):
        w = w.astype(np.float64)
        y = np.zeros((H, W, w.shape[1]))
        for py in range(H):
            for px in range(W):
                acc = b.astype(np.float64).copy()
                for tap in range(9):
                    sy, sx = py + tap // 3 - 1, px + tap % 3 - 1
                    if 0 <= sy < H and 0 <= sx < W:
                        acc += w[tap] @ x[sy, sx]
                y[py, px] = acc
        x = np.maximum(y, 0) if li < len(layers) - 1 else y
    return x


def test_conv_layout_matches_loops():
    rng = np.random.Generator(np.random.PCG64(5))
    feat = np.zeros((4, 5, 16), np.float32)
    feat[..., :8] = rng.uniform(-1, 1, (4, 5, 8))
    p = rand_params(6)
    got, _ = R.forward(p, feat, exact=True)
    want = loop_forward(p, feat)
    np.testing.assert_allclose(got.detach().numpy(), want, rtol=1e-9, atol=1e-12)
    assert not np.allclose(want[..., :3], 0)


def test_single_tap_shift():
    """Weight on tap 5 (dy 0, dx +1) of layer 0 reads the right neighbour."""
    _, n = nnaa.layer_offsets()
    offs, _ = nnaa.layer_offsets()
    p = np.zeros(n, np.float32)
    ci, co, wo, bo = offs[0]
    p[wo + (5 * co + 0) * ci + 0] = 1.0  # layer 0: out ch 0 <- in ch 0 at x + 1
    for li in range(1, 6):
        ci, co, wo, bo = offs[li]
        p[wo + (4 * co + 0) * ci + 0] = 1.0  # centre tap passes channel 0 through
    feat = np.zeros((3, 4, 16), np.float32)
    feat[..., 0] = np.arange(12, dtype=np.float32).reshape(3, 4) + 1
    out, _ = R.forward(p, feat, exact=True)
    o = out.detach().numpy()[..., 0]
    want = np.concatenate([feat[:, 1:, 0], np.zeros((3, 1))], axis=1)
    np.testing.assert_array_equal(o, want)


def test_gradient_matches_central_differences():
    rng = np.random.Generator(np.random.PCG64(9))
    H, W = 4, 5
    color = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    target = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    depth = rng.uniform(0.1, 0.9, (H, W))
    depth[0, 0] = R.MAXF64
    p = rand_params(10, 0.08).astype(np.float64)
    _, g = R.loss_and_grad(p, color, depth, target, exact=True)
    offs, _ = nnaa.layer_offsets()
    idx = [offs[0][2] + 3, offs[2][2] + 100, offs[5][2] + 5, offs[5][3] + 1, offs[3][3] + 7, offs[1][2] + 4000]
    for i in idx:
        h = 1e-5
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = R.loss_and_grad(pp, color, depth, target, exact=True)
        lm, _ = R.loss_and_grad(pm, color, depth, target, exact=True)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-3], np.float32)
    got = R.bf16(x)
    # ties to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 -> 1
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == np.float32(1.0 + 2 ** -6) and got[3] == 1.0
    assert got[4] == -2.5
    np.testing.assert_array_equal(R.bf16(x), R.bf16(torch.tensor(x)).numpy())


def test_features_of_one_pixel():
    color = np.zeros((3, 3, 4), np.uint8)
    color[1, 1] = (255, 0, 51, 255)
    depth = np.full((3, 3), R.MAXF64)
    depth[1, 1] = 0.5
    depth[1, 2] = 1.5
    depth[0, 1] = 0.5
    f = R.features(color, depth)
    # t: 0.5 -> 0, 1.5 -> 1, cleared -> 1
    assert f[1, 1, 0] == 1.0 and f[1, 1, 1] == 0.0 and f[1, 1, 2] == R.bf16(np.float32(0.2))
    assert f[1, 1, 3] == 0.0 and f[1, 2, 3] == 1.0 and f[0, 0, 3] == 1.0
    # gx = t(1,2) - t(1,0) = 1 - 1 = 0; gy = t(2,1) - t(0,1) = 1 - 0 = 1 -> n = (0, -64, 1) / |.|
    n = np.array([0.0, -64.0, 1.0]) / np.sqrt(64.0 ** 2 + 1)
    np.testing.assert_allclose(f[1, 1, 4:7], n, rtol=2 ** -8)
    assert f[1, 1, 7] == 1.0 and f[0, 0, 7] == 0.0
    assert not f[..., 8:].any()
