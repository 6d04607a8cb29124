"""Runs the headline frame a few times for hardware-counter collection, and a
plain device copy of known size first, whose FETCH_SIZE / WRITE_SIZE
readings calibrate the counters (bytes per counter unit).

    rocprofv3 --pmc FETCH_SIZE --kernel-include-regex 'k_|copy' -- \\
        python3 -m fauxgl_amd.profile_frames --frames 3

The tile shape is the autotuner's pick for this workload (or --tile WxH);
the tuning draws run before the frames, under other kernel names
(fgl::t<W>x<H>::...), so counters of the frames are told apart by the shape
printed at the end.

Prints the copy size, the tile shape and the draw statistics of the last
frame.
"""
import argparse
import json
import sys


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=3)
    ap.add_argument("--ssaa", type=int, default=16, choices=[1, 16])
    ap.add_argument("--copy-mib", type=int, default=512)
    ap.add_argument("--tile", default="auto", help="WxH, or auto (the autotuner's pick)")
    args = ap.parse_args()

    import torch  # device memory for the calibration copy only

    from fauxgl_amd import api as A
    from fauxgl_amd import meshes as M

    n = args.copy_mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    copy_bytes = n * 4

    W, H = 1920, 1080
    s = 4 if args.ssaa == 16 else 1
    mesh = M.noise_sphere_mesh()
    big = A.NewContext(W * s, H * s)
    out = A.NewContext(W, H) if s > 1 else None
    big.Shader = M.phong_shader(W, H)
    dm = big.Upload(mesh)
    if args.tile == "auto":
        tile = big.Autotune(dm, reps=3)
    else:
        tile = tuple(int(v) for v in args.tile.split("x"))
        big.SetTileShape(*tile)
    info = None
    for _ in range(args.frames):
        big.ClearColorBufferWith(M.BACKGROUND)
        big.ClearDepthBuffer()
        info = big.DrawMesh(dm)
        if out is not None:
            big.Resolve(out, "tent")
    big.Sync()
    print(json.dumps({"copy_bytes": copy_bytes, "frames": args.frames, "tile": list(tile), "info": list(info),
                      "draw_stats": big.DrawStats()}), flush=True)
    dm.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
