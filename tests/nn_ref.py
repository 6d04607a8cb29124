"""FP32 CPU reference of the learned anti-aliasing network (fauxgl_amd.nnaa),
in numpy + torch on the CPU, for the GPU tests.  BF16 rounding of the
convolution operands is emulated where the GPU rounds (input features,
weights, activations between layers, gradients handed between layers), with
straight-through gradients, so the comparison isolates accumulation order."""
import numpy as np
import torch
import torch.nn.functional as F

from fauxgl_amd import nnaa

MAXF64 = np.finfo(np.float64).max


def bf16(x):
    """Round to nearest even BF16, returned as float32 (numpy or torch)."""
    if isinstance(x, torch.Tensor):
        return x.to(torch.bfloat16).to(torch.float32)
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def features(color, depth):
    """color (H, W, 4) uint8, depth (H, W) float64 -> (H, W, 16) float32."""
    H, W = depth.shape
    drawn = depth != MAXF64
    if drawn.any():
        lo, hi = depth[drawn].min(), depth[drawn].max()
    else:
        lo, hi = MAXF64, -MAXF64
    if hi > lo:
        with np.errstate(over="ignore", invalid="ignore"):
            t = np.where(drawn, ((depth - lo) / (hi - lo)), 1.0).astype(np.float32)
    else:
        t = np.ones((H, W), np.float32)
    xs = np.arange(W)
    ys = np.arange(H)
    xl, xr = np.maximum(xs - 1, 0), np.minimum(xs + 1, W - 1)
    yu, yd = np.maximum(ys - 1, 0), np.minimum(ys + 1, H - 1)
    gx = t[:, xr] - t[:, xl]
    gy = t[yd, :] - t[yu, :]
    sx = np.float32(-64.0) * gx
    sy = np.float32(-64.0) * gy
    rn = np.float32(1.0) / np.sqrt(sx * sx + sy * sy + np.float32(1.0))
    f = np.zeros((H, W, 16), np.float32)
    inv = np.float32(1.0 / 255.0)
    f[..., 0] = color[..., 0].astype(np.float32) * inv
    f[..., 1] = color[..., 1].astype(np.float32) * inv
    f[..., 2] = color[..., 2].astype(np.float32) * inv
    f[..., 3] = t
    f[..., 4] = sx * rn
    f[..., 5] = sy * rn
    f[..., 6] = rn
    f[..., 7] = drawn.astype(np.float32)
    return bf16(f)


class _RoundST(torch.autograd.Function):
    """BF16 rounding forward; gradient rounded to BF16 too (the GPU hands
    BF16 gradients between layers)."""

    @staticmethod
    def forward(ctx, x):
        return bf16(x)

    @staticmethod
    def backward(ctx, g):
        return bf16(g)


class _GradBF16(torch.autograd.Function):
    """Identity forward; the output gradient is stored as BF16."""

    @staticmethod
    def forward(ctx, x):
        return x.clone()

    @staticmethod
    def backward(ctx, g):
        return bf16(g)


def torch_weights(w):
    """[9][Cout][Cin] -> conv2d weight [Cout][Cin][3][3] (tap = 3 * ky + kx)."""
    w = torch.as_tensor(w)
    return w.reshape(3, 3, w.shape[1], w.shape[2]).permute(2, 3, 0, 1).contiguous()


def forward(params, feat, requires_grad=False, exact=False):
    """(H, W, 16) features -> (H, W, 16) raw output, plus the parameter
    tensors (leaves) for autograd.  exact: float64, no BF16 rounding."""
    layers = nnaa.split_params(params)
    dt = torch.float64 if exact else torch.float32
    x = torch.as_tensor(feat).to(dt).permute(2, 0, 1)[None].contiguous()
    leaves = []
    for li, (w, b) in enumerate(layers):
        wt = torch.tensor(np.array(w) if exact else bf16(w), dtype=dt).requires_grad_(requires_grad)
        bt = torch.tensor(np.array(b), dtype=dt).requires_grad_(requires_grad)
        leaves.append((wt, bt))
        y = F.conv2d(x, torch_weights(wt), bt, padding=1)
        if li < len(layers) - 1:
            x = torch.relu(y) if exact else _RoundST.apply(torch.relu(y))
        else:
            x = _GradBF16.apply(y) if (requires_grad and not exact) else y
    return x[0].permute(1, 2, 0), leaves


def loss_and_grad(params, color, depth, target, exact=False):
    """Mean squared error of color/255 + out[..., :3] against target/255 over
    H * W * 3 values, and the gradient of every parameter (flat, the
    network's layout)."""
    feat = features(color, depth)
    out, leaves = forward(params, feat, requires_grad=True, exact=exact)
    dt = torch.float64 if exact else torch.float32
    c = torch.tensor(color[..., :3].astype(np.float32) / np.float32(255.0), dtype=dt)
    t = torch.tensor(target[..., :3].astype(np.float32) / np.float32(255.0), dtype=dt)
    d = c + out[..., :3] - t
    loss = (d * d).mean()
    loss.backward()
    g = []
    for wt, bt in leaves:
        g.append(wt.grad.numpy().reshape(-1))
        g.append(bt.grad.numpy().reshape(-1))
    return float(loss.detach()), np.concatenate(g)
