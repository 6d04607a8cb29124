"""GPU checks of the learned anti-aliasing network (fauxgl_amd.nnaa) against
the FP32 CPU reference in tests/nn_ref.py, which emulates the GPU's BF16
rounding points.  Floating point: the tolerances below are stated per test.
Each case prints its name before touching the GPU."""
import sys

import numpy as np
import pytest

import nn_ref as R
from fauxgl_amd import _native as N
from fauxgl_amd import api as A
from fauxgl_amd import meshes as M
from fauxgl_amd import nnaa
from fauxgl_amd import train_nnaa as T

pytestmark = pytest.mark.gpu


def _say(name):
    print(f"[case] {name}", flush=True)
    sys.stdout.flush()


def random_frame(h, w, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    color = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    color[..., 3] = 255
    depth = rng.uniform(0.2, 0.8, (h, w))
    depth[rng.uniform(size=(h, w)) < 0.2] = R.MAXF64
    return color, depth


def frame_ctx(color, depth):
    h, w = depth.shape
    c = A.NewContext(w, h)
    c.SetColorBuffer(color)
    c.SetDepthBuffer(depth)
    return c


def rand_params(seed, scale=0.08):
    _, n = nnaa.layer_offsets()
    return np.random.Generator(np.random.PCG64(seed)).normal(0, scale, n).astype(np.float32)


@pytest.mark.parametrize("h,w,seed", [(45, 67, 1), (16, 16, 2), (9, 130, 3), (1, 1, 4)])
def test_forward_matches_reference(h, w, seed):
    """Raw output within 5e-3 * max|ref| elementwise and 1e-4 * max on
    average.  The output is exact up to FP32 summation order except where an
    activation lands on the other side of a BF16 rounding boundary (one unit
    in 2^8; a few per layer at this size, each felt over the 13 x 13 pixels
    downstream), which the bounds allow for; a wrong tap, channel or
    operand order gives errors of the order of the output itself."""
    _say(f"nn forward {w}x{h} seed {seed}")
    color, depth = random_frame(h, w, seed)
    ctx = frame_ctx(color, depth)
    net = nnaa.Network(ctx, seed=seed)
    p = rand_params(seed + 100)
    net.SetParams(p)
    got = net.Residual(ctx)
    want = R.forward(p, R.features(color, depth))[0].detach().numpy()
    scale = float(np.abs(want).max()) + 1e-6
    err = np.abs(got - want)
    assert err.max() <= 5e-3 * scale + 1e-6, (err.max(), scale)
    assert err.mean() <= 1e-4 * scale, (err.mean(), scale)
    # the seeded initialisation round-trips through GetParams / SetParams
    net.SetParams(net.GetParams())
    np.testing.assert_array_equal(net.GetParams(), p)


def test_infer_applies_the_residual():
    """Infer's NRGBA output is exactly the clamped, rounded input + residual."""
    _say("nn infer 45x67")
    color, depth = random_frame(45, 67, 7)
    ctx = frame_ctx(color, depth)
    dst = A.NewContext(67, 45)
    net = nnaa.Network(ctx, seed=7)
    net.SetParams(rand_params(8, 0.1))
    res = net.Residual(ctx)
    net.Infer(ctx, dst)
    img = dst.Image()
    v = color[..., :3].astype(np.float32) * np.float32(1.0 / 255.0) + res[..., :3]
    v = np.clip(v, 0.0, 1.0)
    want = (v * np.float32(255.0) + np.float32(0.5)).astype(np.uint32).astype(np.uint8)
    np.testing.assert_array_equal(img[..., :3], want)
    np.testing.assert_array_equal(img[..., 3], color[..., 3])


@pytest.mark.parametrize("h,w,seed", [(37, 53, 11), (8, 16, 12)])
def test_gradient_matches_reference(h, w, seed):
    """Loss within 1e-4 relative; per-layer weight and bias gradients within
    2e-2 relative L2 error (BF16 activations and gradients)."""
    _say(f"nn grad {w}x{h} seed {seed}")
    color, depth = random_frame(h, w, seed)
    target, _ = random_frame(h, w, seed + 1)
    src = frame_ctx(color, depth)
    tgt = A.NewContext(w, h)
    tgt.SetColorBuffer(target)
    net = nnaa.Network(src, seed=seed)
    p = rand_params(seed + 200)
    net.SetParams(p)
    loss, g = net.Grad(src, tgt)
    rloss, rg = R.loss_and_grad(p, color, depth, target)
    assert abs(loss - rloss) <= 1e-4 * abs(rloss), (loss, rloss)
    offs, _ = nnaa.layer_offsets()
    for li, (ci, co, wo, bo) in enumerate(offs):
        for name, a, b in (("w", wo, wo + 9 * ci * co), ("b", bo, bo + co)):
            gg, rr = g[a:b], rg[a:b]
            nrm = np.linalg.norm(rr)
            if nrm == 0:
                assert np.abs(gg).max() == 0, (li, name)
                continue
            rel = np.linalg.norm(gg - rr) / nrm
            assert rel <= 2e-2, (li, name, rel)


def test_size_mismatch_is_an_error():
    _say("nn size mismatch")
    a = A.NewContext(32, 16)
    b = A.NewContext(16, 16)
    net = nnaa.Network(a)
    with pytest.raises(N.FglError):
        net.Infer(a, b)
    with pytest.raises(N.FglError):
        net.Grad(a, b)
    with pytest.raises(ValueError):
        net.SetParams(np.zeros(3, np.float32))


def test_training_beats_the_1x_frame():
    """150 Adam steps on a 192x108 sphere-and-cube scene lower the loss and
    bring the network's output closer to the 16x frame than the 1x frame is,
    on a view not trained on."""
    _say("nn training 192x108")
    pr = T.PairRenderer(192, 108, M.sphere_and_cube_mesh())
    net = nnaa.Network(pr.lo, seed=3)
    rng = np.random.Generator(np.random.PCG64(5))
    losses = []
    for s in range(150):
        pr.render(T.random_eye(rng))
        losses.append(net.TrainStep(pr.lo, pr.hi, 2e-3))
    assert np.mean(losses[-20:]) < 0.8 * np.mean(losses[:5]), (losses[:5], losses[-20:])
    (p1, pn), = T.evaluate(net, pr, [T.HEADLINE_EYE])
    print(f"psnr 1x {p1:.2f} dB, network {pn:.2f} dB", flush=True)
    assert pn > p1
