"""ctypes binding of libfauxgl_amd.so, the gfx950 draw path (include/fgl.h).

There is no CPU fallback: if the library is missing or cannot be loaded,
every entry point raises.  Build it with ``python -m fauxgl_amd.build``.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# FAUXGL_AMD_LIB selects another build of the same library (e.g. one built
# with different tile constants); the default is the in-tree build.
LIB_PATH = os.environ.get("FAUXGL_AMD_LIB") or os.path.join(HERE, "libfauxgl_amd.so")

FGL_OK = 0
FGL_E_INVALID = -1
FGL_E_NOMEM = -2
FGL_E_DEVICE = -3
FGL_E_UNSUPPORTED = -4
FGL_E_IO = -5

FGL_FACE_CW, FGL_FACE_CCW = 1, 2
FGL_CULL_NONE, FGL_CULL_FRONT, FGL_CULL_BACK = 1, 2, 3
FGL_SHADER_SOLID, FGL_SHADER_TEXTURE, FGL_SHADER_PHONG = 1, 2, 3
FGL_TEX_NRGBA8, FGL_TEX_RGBA8, FGL_TEX_RGBA16 = 1, 2, 3
FGL_RESOLVE_TENT, FGL_RESOLVE_BOX = 1, 2

# Every symbol include/fgl.h declares.
EXPORTED = [
    "fgl_init", "fgl_last_error", "fgl_ctx_create", "fgl_ctx_destroy", "fgl_ctx_size", "fgl_clear_color",
    "fgl_clear_depth", "fgl_texture_create", "fgl_texture_destroy", "fgl_draw_triangles", "fgl_draw_lines",
    "fgl_mesh_create", "fgl_mesh_destroy", "fgl_draw_mesh", "fgl_read_color", "fgl_read_depth",
    "fgl_write_color", "fgl_write_depth", "fgl_depth_image", "fgl_depth_image_noalloc_async", "fgl_resolve", "fgl_profile_enable",
    "fgl_profile_read", "fgl_profile_reset", "fgl_ctx_sync", "fgl_draw_stats", "fgl_profile_only",
    "fgl_nn_create", "fgl_nn_destroy", "fgl_nn_param_count", "fgl_nn_get_params", "fgl_nn_set_params",
    "fgl_nn_infer", "fgl_nn_residual", "fgl_nn_grad", "fgl_nn_train_step",
    "fgl_mesh_counts", "fgl_mesh_read", "fgl_mesh_bounding_box", "fgl_mesh_transform", "fgl_mesh_fit_inside",
    "fgl_mesh_smooth_normals", "fgl_mesh_smooth_normals_threshold", "fgl_mesh_load_stl", "fgl_mesh_load_obj",
    "fgl_encode_png_bound", "fgl_encode_png",
    "fgl_batch_create", "fgl_batch_destroy", "fgl_batch_error", "fgl_batch_add", "fgl_batch_run",
    "fgl_tile_shapes", "fgl_ctx_set_tile", "fgl_ctx_get_tile", "fgl_autotune_mesh", "fgl_tune_save",
    "fgl_tune_load", "fgl_tune_clear", "fgl_ctx_force_capacities",
]

_d = ctypes.c_double


class FglDrawStat(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("total_pixels", "updated_pixels", "clipped_triangles",
                                                "side_triangles", "pairs", "spans", "fragments",
                                                "overflow_mask", "reruns")]


class FglState(ctypes.Structure):
    _fields_ = [("read_depth", ctypes.c_uint8), ("write_depth", ctypes.c_uint8), ("write_color", ctypes.c_uint8),
                ("alpha_blend", ctypes.c_uint8), ("wireframe", ctypes.c_uint8), ("pad", ctypes.c_uint8 * 3),
                ("front_face", ctypes.c_int32), ("cull", ctypes.c_int32), ("line_width", _d),
                ("depth_bias", _d)]


class FglShader(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("pad", ctypes.c_int32), ("matrix", _d * 16), ("color", _d * 4),
                ("light_dir", _d * 3), ("camera_pos", _d * 3), ("object_color", _d * 4), ("ambient", _d * 4),
                ("diffuse", _d * 4), ("specular", _d * 4), ("specular_power", _d), ("texture", ctypes.c_int64)]


class FglPrims(ctypes.Structure):
    _fields_ = [("count", ctypes.c_int64), ("position", ctypes.POINTER(_d)), ("normal", ctypes.POINTER(_d)),
                ("texture", ctypes.POINTER(_d)), ("color", ctypes.POINTER(_d))]


class FglKernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("launches", ctypes.c_uint64), ("total_ms", _d)]


class FglError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fgl error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Loads libfauxgl_amd.so; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m fauxgl_amd.build` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    vpp = ctypes.POINTER(ctypes.c_void_p)
    i32 = ctypes.c_int32
    u64p = ctypes.POINTER(ctypes.c_uint64)
    sigs = {
        "fgl_init": (i32, [i32]),
        "fgl_last_error": (ctypes.c_char_p, [vp]),
        "fgl_ctx_create": (i32, [i32, i32, vpp]),
        "fgl_ctx_destroy": (None, [vp]),
        "fgl_ctx_size": (i32, [vp, ctypes.POINTER(i32), ctypes.POINTER(i32)]),
        "fgl_clear_color": (i32, [vp, ctypes.POINTER(ctypes.c_uint8)]),
        "fgl_clear_depth": (i32, [vp, _d]),
        "fgl_texture_create": (i32, [vp, i32, i32, i32, vp, ctypes.POINTER(ctypes.c_int64)]),
        "fgl_texture_destroy": (i32, [vp, ctypes.c_int64]),
        "fgl_draw_triangles": (i32, [vp, ctypes.POINTER(FglState), ctypes.POINTER(FglShader),
                                     ctypes.POINTER(FglPrims), u64p]),
        "fgl_draw_lines": (i32, [vp, ctypes.POINTER(FglState), ctypes.POINTER(FglShader),
                                 ctypes.POINTER(FglPrims), u64p]),
        "fgl_mesh_create": (i32, [vp, ctypes.POINTER(FglPrims), ctypes.POINTER(FglPrims), vpp]),
        "fgl_mesh_destroy": (None, [vp]),
        "fgl_draw_mesh": (i32, [vp, ctypes.POINTER(FglState), ctypes.POINTER(FglShader), vp, u64p]),
        "fgl_read_color": (i32, [vp, ctypes.POINTER(ctypes.c_uint8)]),
        "fgl_read_depth": (i32, [vp, ctypes.POINTER(_d)]),
        "fgl_write_color": (i32, [vp, ctypes.POINTER(ctypes.c_uint8)]),
        "fgl_write_depth": (i32, [vp, ctypes.POINTER(_d)]),
        "fgl_depth_image": (i32, [vp, ctypes.POINTER(ctypes.c_uint16)]),
        "fgl_depth_image_noalloc_async": (i32, [vp, vp]),
        "fgl_resolve": (i32, [vp, vp, i32]),
        "fgl_profile_enable": (i32, [vp, i32]),
        "fgl_profile_read": (i32, [vp, ctypes.POINTER(FglKernelTime), i32, ctypes.POINTER(i32)]),
        "fgl_profile_reset": (i32, [vp]),
        "fgl_ctx_sync": (i32, [vp]),
        "fgl_profile_only": (i32, [vp, ctypes.c_char_p]),
        "fgl_draw_stats": (i32, [vp, ctypes.POINTER(FglDrawStat)]),
        "fgl_ctx_force_capacities": (i32, [vp, ctypes.POINTER(ctypes.c_uint32)]),
        "fgl_mesh_counts": (i32, [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
        "fgl_mesh_read": (i32, [vp, i32, i32, ctypes.POINTER(_d), ctypes.c_int64]),
        "fgl_mesh_bounding_box": (i32, [vp, ctypes.POINTER(_d)]),
        "fgl_mesh_transform": (i32, [vp, ctypes.POINTER(_d)]),
        "fgl_mesh_fit_inside": (i32, [vp, ctypes.POINTER(_d), ctypes.POINTER(_d), ctypes.POINTER(_d)]),
        "fgl_mesh_smooth_normals": (i32, [vp]),
        "fgl_mesh_smooth_normals_threshold": (i32, [vp, _d]),
        "fgl_mesh_load_stl": (i32, [vp, ctypes.POINTER(ctypes.c_uint8), ctypes.c_int64, vpp]),
        "fgl_mesh_load_obj": (i32, [vp, ctypes.c_char_p, ctypes.c_int64, vpp]),
        "fgl_batch_create": (i32, [vpp]),
        "fgl_batch_destroy": (None, [vp]),
        "fgl_batch_error": (ctypes.c_char_p, [vp]),
        "fgl_batch_add": (i32, [vp, vp, ctypes.POINTER(ctypes.c_uint8), _d, ctypes.POINTER(FglState),
                                ctypes.POINTER(FglShader), vp]),
        "fgl_batch_run": (i32, [vp, u64p]),
        "fgl_tile_shapes": (i32, [ctypes.POINTER(i32), i32, ctypes.POINTER(i32)]),
        "fgl_ctx_set_tile": (i32, [vp, i32, i32]),
        "fgl_ctx_get_tile": (i32, [vp, ctypes.c_int64, ctypes.POINTER(i32)]),
        "fgl_autotune_mesh": (i32, [vp, ctypes.POINTER(FglState), ctypes.POINTER(FglShader), vp, i32,
                                    ctypes.POINTER(i32)]),
        "fgl_tune_save": (i32, [ctypes.c_char_p]),
        "fgl_tune_load": (i32, [ctypes.c_char_p]),
        "fgl_tune_clear": (i32, []),
        "fgl_encode_png_bound": (i32, [vp, ctypes.POINTER(ctypes.c_int64)]),
        "fgl_encode_png": (i32, [vp, ctypes.POINTER(ctypes.c_uint8), ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc, ctx=None):
    if rc != FGL_OK:
        msg = load().fgl_last_error(ctx)
        raise FglError(rc, msg.decode() if msg else "")
    return rc
