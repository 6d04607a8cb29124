"""Learned anti-aliasing on the GPU (libfauxgl_amd.so, fgl_nn_* entry points).

A six-layer 3x3 convolutional network (BF16 operands on the matrix cores,
FP32 accumulation) maps a 1x render's color, depth and depth-derived normals
to an estimate of the same frame rendered with 4x4 supersampling and a tent
resolve.  It has no counterpart in fauxgl: it replaces the 16x-sized draw and
the resolve (``Context.Resolve``) with a 1x draw and one network pass.

Parameters are FP32, per layer weights [9][Cout][Cin] (tap = 3 * (dy + 1)
+ (dx + 1), cross-correlation like ``torch.nn.functional.conv2d``) then bias
[Cout].  Input features (16 channels, 8..15 zero): R, G, B / 255, depth t in
[0, 1] (``DepthImage``'s normalisation, 1 where cleared), the unit normal of
the surface (-64 dt/dx, -64 dt/dy, 1) from central differences of t, and the
coverage mask.  The prediction is the input color plus output channels 0..2.
"""
import ctypes
import os

import numpy as np

from . import _native as N

LAYERS = [(16, 32), (32, 32), (32, 32), (32, 32), (32, 32), (32, 16)]
FEAT_C = 16
OUT_C = 16
HERE = os.path.dirname(os.path.abspath(__file__))
# weights trained on engine renders (fauxgl_amd/train_nnaa.py)
DEFAULT_WEIGHTS = os.path.join(HERE, "nnaa_weights.npy")


def layer_offsets():
    """[(cin, cout, weight offset, bias offset)] into the flat parameters."""
    out, off = [], 0
    for ci, co in LAYERS:
        w = off
        off += 9 * ci * co
        out.append((ci, co, w, off))
        off += co
    return out, off


def split_params(p):
    """Flat parameters -> [(weights [9][Cout][Cin], bias [Cout])]."""
    offs, n = layer_offsets()
    p = np.asarray(p)
    if p.dtype != np.float64:
        p = p.astype(np.float32)
    if p.size != n:
        raise ValueError(f"expected {n} parameters, got {p.size}")
    return [(p[w:w + 9 * ci * co].reshape(9, co, ci), p[b:b + co]) for ci, co, w, b in offs]


class Network:
    """The network's parameters and optimiser state on ctx's device."""

    def __init__(self, ctx, seed=0, weights=None):
        L = N.load()
        self._L = L
        self._sigs(L)
        h = ctypes.c_void_p()
        N.check(L.fgl_nn_create(ctx._ptr(), ctypes.c_uint64(seed), ctypes.byref(h)), ctx._ptr())
        self.h = h
        self.ctx = ctx
        n = ctypes.c_int64()
        N.check(L.fgl_nn_param_count(self.h, ctypes.byref(n)))
        self.n = int(n.value)
        if weights is not None:
            self.SetParams(np.load(weights) if isinstance(weights, str) else weights)

    @staticmethod
    def _sigs(L):
        if getattr(L, "_nn_sigs", False):
            return
        vp, i32, i64, f32p = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(ctypes.c_float)
        sigs = {
            "fgl_nn_create": (i32, [vp, ctypes.c_uint64, ctypes.POINTER(vp)]),
            "fgl_nn_destroy": (None, [vp]),
            "fgl_nn_param_count": (i32, [vp, ctypes.POINTER(i64)]),
            "fgl_nn_get_params": (i32, [vp, f32p, i64]),
            "fgl_nn_set_params": (i32, [vp, f32p, i64]),
            "fgl_nn_infer": (i32, [vp, vp, vp]),
            "fgl_nn_residual": (i32, [vp, vp, f32p, i64]),
            "fgl_nn_grad": (i32, [vp, vp, vp, f32p, f32p, i64]),
            "fgl_nn_train_step": (i32, [vp, vp, vp, ctypes.c_float, f32p]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L._nn_sigs = True

    def _ptr(self):
        if not self.h:
            raise RuntimeError("network is closed")
        return self.h

    def close(self):
        if getattr(self, "h", None):
            self._L.fgl_nn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def GetParams(self):
        p = np.empty(self.n, dtype=np.float32)
        N.check(self._L.fgl_nn_get_params(self._ptr(), p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), self.n),
                self.ctx._ptr())
        return p

    def SetParams(self, p):
        """Loads parameters and resets the optimiser."""
        p = np.ascontiguousarray(np.asarray(p, dtype=np.float32).reshape(-1))
        if p.size != self.n:
            raise ValueError(f"expected {self.n} parameters, got {p.size}")
        N.check(self._L.fgl_nn_set_params(self._ptr(), p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), self.n),
                self.ctx._ptr())

    def Infer(self, src, dst):
        """dst's color = the network's estimate of src's supersampled frame."""
        N.check(self._L.fgl_nn_infer(self._ptr(), src._ptr(), dst._ptr()), src._ptr())

    def Residual(self, src):
        """Raw output for src, (H, W, 16) float32 (channels 0..2 are used)."""
        out = np.empty((src.Height, src.Width, OUT_C), dtype=np.float32)
        N.check(self._L.fgl_nn_residual(self._ptr(), src._ptr(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                        out.size), src._ptr())
        return out

    def Grad(self, src, target):
        """(loss, gradient of every parameter) for the pair src -> target."""
        loss = ctypes.c_float()
        g = np.empty(self.n, dtype=np.float32)
        N.check(self._L.fgl_nn_grad(self._ptr(), src._ptr(), target._ptr(), ctypes.byref(loss),
                                    g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), self.n), src._ptr())
        return float(loss.value), g

    def TrainStep(self, src, target, lr=1e-3):
        """One Adam step; returns the loss before the step."""
        loss = ctypes.c_float()
        N.check(self._L.fgl_nn_train_step(self._ptr(), src._ptr(), target._ptr(), float(lr), ctypes.byref(loss)),
                src._ptr())
        return float(loss.value)

    def Save(self, path):
        np.save(path, self.GetParams())
