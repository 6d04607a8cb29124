"""Trains the learned anti-aliasing network on frames the draw path renders.

Each step renders one view at 1x (the network's input) and the same view at
4x4 the size followed by the tent resolve (the target), then takes one Adam
step on the GPU.  Views orbit the scene at random azimuth, elevation and
distance, with a random mesh (the headline noise sphere, two other noise
spheres, the sphere-and-cube scene), light direction, object color and
background; the evaluation views (the headline camera among them, with the
headline light and colors) are never trained on.

    python -m fauxgl_amd.train_nnaa --steps 1500 --out fauxgl_amd/nnaa_weights.npy
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import api as A
from . import meshes as M
from . import nnaa


def psnr(a, b):
    """PSNR over R, G, B of two (H, W, 4) uint8 images, in dB."""
    d = a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)
    mse = float(np.mean(d * d))
    return float("inf") if mse == 0 else 10.0 * math.log10(255.0 * 255.0 / mse)


def orbit_eye(azimuth, elevation, distance):
    c = M.CENTER
    return A.V(c.X + distance * math.cos(elevation) * math.cos(azimuth), c.Y + distance * math.sin(elevation),
               c.Z + distance * math.cos(elevation) * math.sin(azimuth))


HEADLINE_EYE = M.EYE


class PairRenderer:
    """Renders (1x, 16x-resolved) pairs of W x H frames of one or more
    meshes resident on the GPU."""

    def __init__(self, width, height, meshes):
        self.W, self.H = width, height
        self.lo = A.NewContext(width, height)
        self.big = A.NewContext(4 * width, 4 * height)
        self.hi = A.NewContext(width, height)
        self.out = A.NewContext(width, height)
        if not isinstance(meshes, (list, tuple)):
            meshes = [meshes]
        self.dms = [self.big.Upload(m) for m in meshes]
        self.dm = self.dms[0]

    def render(self, eye, mesh=0, light=None, color=None, background=None):
        """The frame of mesh index `mesh` seen from eye (headline light,
        object color and background unless given)."""
        for ctx, w, h in ((self.lo, self.W, self.H), (self.big, 4 * self.W, 4 * self.H)):
            ctx.ClearColorBufferWith(background if background is not None else M.BACKGROUND)
            ctx.ClearDepthBuffer()
            sh = M.phong_shader(w, h, eye=eye)
            if light is not None:
                sh.LightDirection = A.Vector(*light)
            if color is not None:
                sh.ObjectColor = A.Color(*color)
            ctx.Shader = sh
            ctx.DrawMesh(self.dms[mesh])
        self.big.Resolve(self.hi, "tent")

    def close(self):
        for dm in self.dms:
            dm.close()
        for c in (self.lo, self.big, self.hi, self.out):
            c.close()


def training_meshes():
    """The headline noise sphere, two smaller noise spheres of other seeds
    and amplitudes, and the sphere-and-cube scene."""
    return [M.noise_sphere_mesh(),
            M.noise_sphere_mesh(ntri=300000, detail=7, amplitude=0.25, seed=7),
            M.noise_sphere_mesh(ntri=80000, detail=6, amplitude=0.08, seed=11),
            M.sphere_and_cube_mesh()]


def random_scene(rng, nmesh):
    """(mesh index, light direction, object color, background) drawn at random."""
    mesh = int(rng.integers(nmesh))
    v = rng.normal(size=3)
    v[1] = abs(v[1]) + 0.2
    light = tuple(v / np.linalg.norm(v))
    color = tuple(rng.uniform(0.1, 0.9, 3)) + (1.0,)
    bg = tuple(rng.uniform(0.0, 1.0, 3)) + (1.0,) if rng.uniform() < 0.5 else M.BACKGROUND
    return mesh, light, color, bg


def random_eye(rng):
    az = rng.uniform(0, 2 * math.pi)
    el = rng.uniform(-0.6, 0.9)
    dist = rng.uniform(2.6, 4.5)
    return orbit_eye(az, el, dist)


def eval_eyes():
    """The headline camera plus three fixed held-out views."""
    return [HEADLINE_EYE, orbit_eye(0.9, 0.35, 3.4), orbit_eye(2.8, -0.2, 3.0), orbit_eye(4.4, 0.6, 4.0)]


def evaluate(net, pr, eyes):
    """[(psnr of 1x vs target, psnr of network output vs target)] per view."""
    res = []
    for e in eyes:
        pr.render(e)
        net.Infer(pr.lo, pr.out)
        hi = pr.hi.Image()
        res.append((psnr(pr.lo.Image(), hi), psnr(pr.out.Image(), hi)))
    return res


def train(steps=1500, width=1920, height=1080, lr=1e-3, seed=0, init=None, log=None, meshes=None):
    rng = np.random.Generator(np.random.PCG64(seed))
    pr = PairRenderer(width, height, meshes if meshes is not None else training_meshes())
    net = nnaa.Network(pr.lo, seed=seed, weights=init)
    t0 = time.time()
    losses = []
    for s in range(steps):
        mesh, light, color, bg = random_scene(rng, len(pr.dms))
        pr.render(random_eye(rng), mesh, light, color, bg)
        frac = s / max(steps - 1, 1)
        cur = lr * (0.5 * (1 + math.cos(math.pi * frac)) * 0.95 + 0.05)
        losses.append(net.TrainStep(pr.lo, pr.hi, cur))
        if log and (s % 50 == 0 or s == steps - 1):
            log(f"step {s} loss {np.mean(losses[-50:]):.6g} lr {cur:.3g} {time.time() - t0:.1f}s")
    return net, pr, losses


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None)
    ap.add_argument("--out", default=nnaa.DEFAULT_WEIGHTS)
    args = ap.parse_args()

    def log(m):
        print(m, flush=True)

    net, pr, losses = train(args.steps, args.width, args.height, args.lr, args.seed, args.init, log)
    net.Save(args.out)
    ev = evaluate(net, pr, eval_eyes())
    print(json.dumps({"weights": args.out, "steps": args.steps, "final_loss": float(np.mean(losses[-50:])),
                      "eval_psnr_1x_vs_16x": [round(a, 3) for a, _ in ev],
                      "eval_psnr_nn_vs_16x": [round(b, 3) for _, b in ev]}), flush=True)
    net.close()
    pr.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
