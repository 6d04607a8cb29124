"""The batched multi-view and neural anti-aliasing workloads at their full
sizes, as bench.py runs them.

- Batched views: 32 noise spheres of 1k-1M triangles, 256 views at 256^2,
  512^2, 1024^2 and 1920x1080, meshes prepared on the device (BiUnitCube,
  SmoothNormalsThreshold 30 degrees), tile shapes from the autotuner.  Every
  view of every run (the eager run that records the graph, then two graph
  replays) equals the same view drawn on its own, bit for bit; a sample of
  views covering every resolution and the smallest and largest mesh equals
  the oracle (reference: context.go:435-439, DrawMesh per view).
- Neural anti-aliasing: the committed network on the 1920x1080 headline 1x
  frame, against the FP32 CPU reference with BF16 rounding emulated
  (tests/nn_ref.py), within the tolerances of test_gpu_nn.py; and the
  network's output is closer to the 16x SSAA frame than the 1x frame is.

Each case prints its name before touching the GPU."""
import hashlib
import math
import sys

import numpy as np
import pytest

import nn_ref as R
from fauxgl_amd import api as A
from fauxgl_amd import meshes as M
from fauxgl_amd import nnaa
from helpers import compare, run_oracle

pytestmark = pytest.mark.gpu


def _say(name):
    print(f"[case] {name}", flush=True)
    sys.stdout.flush()


def frame_digest(ctx):
    h = hashlib.sha1()
    h.update(ctx.Image().tobytes())
    h.update(ctx.DepthBuffer.tobytes())
    return h.hexdigest()


@pytest.mark.timeout(600)
def test_batched_views_full_workload():
    _say("batched views: 32 meshes, 256 views (workload generation)")
    meshes, views = M.batch_workload()
    assert len(meshes) == 32 and len(views) == 256
    home = A.NewContext(8, 8)
    dms = []
    for p in meshes:
        dm = home.Upload(A.Mesh(p))
        dm.BiUnitCube()
        dm.SmoothNormalsThreshold(math.radians(30.0))
        dms.append(dm)
    ctxs = []
    for m, eye, w, h in views:
        c = A.NewContext(w, h)
        c.Shader = M.batch_shader(w, h, eye)
        ctxs.append(c)
    _say("batched views: autotune")
    tuned = {}
    for (m, eye, w, h), c in zip(views, ctxs):
        key = (w, h, len(meshes[m].position).bit_length())
        if key not in tuned:
            tuned[key] = c.Autotune(dms[m], reps=1)
    assert len(set(tuned.values())) >= 2
    # each view drawn on its own (one context per resolution, reused)
    _say("batched views: single draws")
    solo = {}
    want = []
    for m, eye, w, h in views:
        c = solo.setdefault((w, h), A.NewContext(w, h))
        c.ClearColorBufferWith(M.BACKGROUND)
        c.ClearDepthBuffer()
        c.Shader = M.batch_shader(w, h, eye)
        info = c.DrawMesh(dms[m])
        want.append((info, frame_digest(c)))
    b = A.Batch()
    for (m, eye, w, h), c in zip(views, ctxs):
        b.Add(c, dms[m], clear_color=M.BACKGROUND)
    for run in range(3):  # eager run + recording, then two replays
        _say(f"batched views: run {run}")
        infos = b.Run()
        for i, c in enumerate(ctxs):
            assert infos[i] == want[i][0], (run, i)
            assert frame_digest(c) == want[i][1], (run, i)
    # the oracle on a sample: two views per resolution (smallest and largest
    # mesh there), the smallest and largest mesh overall, and four more
    sizes = [len(meshes[m].position) for m, _, _, _ in views]
    pick = set()
    for res in M.BATCH_RESOLUTIONS:
        idx = [i for i, v in enumerate(views) if (v[2], v[3]) == res]
        pick.add(min(idx, key=lambda i: sizes[i]))
        pick.add(max(idx, key=lambda i: sizes[i]))
    pick.add(int(np.argmin(sizes)))
    pick.add(int(np.argmax(sizes)))
    rng = np.random.Generator(np.random.PCG64(3))
    while len(pick) < 14:
        pick.add(int(rng.integers(len(views))))
    host = {}
    for i in sorted(pick):
        m, eye, w, h = views[i]
        _say(f"batched views: oracle, view {i} ({sizes[i]} triangles, {w}x{h})")
        if m not in host:
            host[m] = dms[m].Download()
        ref = run_oracle(w, h, [("clear_color", M.BACKGROUND), ("shader", M.batch_shader(w, h, eye)),
                                ("mesh", host[m])])
        got = (ctxs[i].Image(), ctxs[i].DepthBuffer, [tuple(infos[i])])
        res = compare(ref, got, color_tol=0)
        assert res["ok"], (i, res)
    b.close()


@pytest.mark.timeout(600)
def test_neural_aa_on_the_headline_frame():
    W, H = 1920, 1080
    _say("neural AA: 1x and 16x headline frames")
    mesh = M.noise_sphere_mesh()
    lo = A.NewContext(W, H)
    lo.Shader = M.phong_shader(W, H)
    dm = lo.Upload(mesh)
    lo.ClearColorBufferWith(M.BACKGROUND)
    lo.ClearDepthBuffer()
    lo.DrawMesh(dm)
    color, depth = lo.Image(), lo.DepthBuffer
    big = A.NewContext(4 * W, 4 * H)
    big.Shader = M.phong_shader(W, H)
    big.ClearColorBufferWith(M.BACKGROUND)
    big.ClearDepthBuffer()
    big.DrawMesh(dm)
    ref16 = A.NewContext(W, H)
    big.Resolve(ref16, "tent")
    img16 = ref16.Image()
    _say("neural AA: network on the 1x frame")
    net = nnaa.Network(lo, weights=nnaa.DEFAULT_WEIGHTS)
    got = net.Residual(lo)
    out = A.NewContext(W, H)
    net.Infer(lo, out)
    img = out.Image()
    _say("neural AA: CPU reference")
    import torch
    torch.set_num_threads(16)
    with torch.no_grad():
        want = R.forward(net.GetParams(), R.features(color, depth))[0].numpy()
    scale = float(np.abs(want).max()) + 1e-6
    err = np.abs(got - want)
    assert err.max() <= 5e-3 * scale + 1e-6, (err.max(), scale)
    assert err.mean() <= 1e-4 * scale, (err.mean(), scale)

    def psnr(a, b):
        d = a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)
        return 10.0 * np.log10(255.0 ** 2 / float(np.mean(d * d)))

    p_net, p_1x = psnr(img, img16), psnr(color, img16)
    print(f"[nn] PSNR vs 16x: network {p_net:.2f} dB, plain 1x {p_1x:.2f} dB", flush=True)
    assert p_net > p_1x and p_net >= 40.0
    net.close()
