"""ctypes mirror of include/aurora.h (the libaurora C ABI).

Struct layouts here must match the header field for field; tests check the
sizes against the C compiler's (tests/test_abi.py).
"""
import ctypes as C
import os

ABI_VERSION = 6  # AURORA_ABI_VERSION of the header this mirror follows

AURORA_OK = 0
AURORA_ERR_INVALID = 1
AURORA_ERR_RUNTIME = 2
AURORA_ERR_NOMEM = 3
AURORA_ERR_STATE = 4

QUAD_PARALLELOGRAM = 0
QUAD_TRIANGLE = 1

MAT_LAMBERTIAN = 0
MAT_METAL = 1
MAT_DIELECTRIC = 2
MAT_DIFFUSE_LIGHT = 3

BVH4_EMPTY = 0xFFFFFFFF
BVH4_PAIRWISE = 0
BVH4_GREEDY = 1
TRAVERSAL_STACK = 96  # entries of the device traversal stack
NOALLOC_MAX_RAYS = 262144  # AURORA_NOALLOC_MAX_RAYS
NOALLOC_SCRATCH_BYTES = 52430848  # AURORA_NOALLOC_SCRATCH_BYTES
TEX_SOLID = 0
TEX_CHECKER = 1
TEX_NOISE = 2
TEX_IMAGE = 3

RENDER_AOV = 1
RENDER_COUNT_FETCHES = 2
RENDER_GRAPH = 4
LOSS_LOG_L2 = 0
LOSS_REL_L2 = 1

TRANSFER_GAMMA2 = 0
TRANSFER_GAMMA22 = 1
TRANSFER_LINEAR = 2

NODE_LEAF = 0x80000000
MISS = -1


class Material(C.Structure):
    _fields_ = [("type", C.c_uint32), ("texture", C.c_uint32), ("fuzz", C.c_float), ("ior", C.c_float)]


class Texture(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("even", C.c_uint32),
        ("odd", C.c_uint32),
        ("seed", C.c_uint32),
        ("color", C.c_float * 3),
        ("scale", C.c_float),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("texel_offset", C.c_uint64),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("n_spheres", C.c_uint32),
        ("spheres", C.c_void_p),
        ("sphere_material", C.c_void_p),
        ("n_quads", C.c_uint32),
        ("quads", C.c_void_p),
        ("quad_material", C.c_void_p),
        ("quad_kind", C.c_void_p),
        ("n_materials", C.c_uint32),
        ("materials", C.c_void_p),
        ("n_textures", C.c_uint32),
        ("textures", C.c_void_p),
        ("n_texels", C.c_uint64),
        ("texels", C.c_void_p),
    ]


class BvhNode(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("child", C.c_uint32), ("hi", C.c_float * 3), ("parent", C.c_uint32)]


class BvhStats(C.Structure):
    _fields_ = [
        ("n_prims", C.c_uint32),
        ("n_nodes", C.c_uint32),
        ("depth", C.c_uint32),
        ("n_wide", C.c_uint32),
        ("wide_depth", C.c_uint32),
        ("wide_form", C.c_uint32),
        ("scene_lo", C.c_float * 3),
        ("scene_hi", C.c_float * 3),
        ("centroid_lo", C.c_float * 3),
        ("centroid_hi", C.c_float * 3),
        ("build_ms", C.c_double),
    ]


class TraceStats(C.Structure):
    _fields_ = [("extend_ms", C.c_double), ("nodes_fetched", C.c_uint64), ("spheres_fetched", C.c_uint64),
                ("quads_fetched", C.c_uint64)]


class CameraDesc(C.Structure):
    _fields_ = [
        ("image_width", C.c_uint32),
        ("image_height", C.c_uint32),
        ("vfov", C.c_float),
        ("lookfrom", C.c_float * 3),
        ("lookat", C.c_float * 3),
        ("vup", C.c_float * 3),
        ("defocus_angle", C.c_float),
        ("focus_dist", C.c_float),
    ]


class RenderDesc(C.Structure):
    _fields_ = [
        ("spp", C.c_uint32),
        ("sample_offset", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("seed", C.c_uint32),
        ("background", C.c_float * 3),
        ("flags", C.c_uint32),
        ("spp_per_pass", C.c_uint32),
        ("tmin", C.c_float),
    ]


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", CameraDesc), ("render", RenderDesc), ("d_rgb_sum", C.c_void_p),
                ("d_albedo_sum", C.c_void_p), ("d_normal_sum", C.c_void_p)]


class RenderStats(C.Structure):
    _fields_ = [
        ("paths", C.c_uint64),
        ("segments", C.c_uint64),
        ("kernel_ms", C.c_double),
        ("extend_ms", C.c_double),
        ("passes", C.c_uint32),
        ("max_live", C.c_uint32),
        ("extend_launches", C.c_uint32),
        ("tail_launches", C.c_uint32),
        ("nodes_fetched", C.c_uint64),
        ("spheres_fetched", C.c_uint64),
        ("quads_fetched", C.c_uint64),
        ("tail_ms", C.c_double),
    ]


PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(PKG_DIR), "libaurora.so")

# (name, restype, argtypes) of every entry point declared in include/aurora.h
_vp = C.c_void_p
_u32 = C.c_uint32
_f = C.c_float
EXPORTS = [
    ("aurora_abi_version", C.c_int, []),
    ("aurora_create", C.c_int, [C.c_int, C.POINTER(_vp)]),
    ("aurora_destroy", None, [_vp]),
    ("aurora_last_error", C.c_char_p, [_vp]),
    ("aurora_synchronize", C.c_int, [_vp]),
    ("aurora_stream", _vp, [_vp]),
    ("aurora_set_launch", C.c_int, [_vp, C.c_int, C.c_int, _u32]),
    ("aurora_set_ray_sort", C.c_int, [_vp, C.c_int]),
    ("aurora_set_render_tree", C.c_int, [_vp, C.c_int]),
    ("aurora_set_queue", C.c_int, [_vp, _u32, _u32]),
    ("aurora_scene_upload", C.c_int, [_vp, C.POINTER(SceneDesc), C.POINTER(_vp)]),
    ("aurora_scene_destroy", None, [_vp]),
    ("aurora_bvh_build", C.c_int, [_vp, C.POINTER(BvhStats)]),
    ("aurora_bvh_download", C.c_int, [_vp, _vp, _u32]),
    ("aurora_bvh_download_wide", C.c_int, [_vp, _vp, _u32]),
    ("aurora_bvh_download_order", C.c_int, [_vp, _vp, _vp, _u32]),
    ("aurora_trace_closest", C.c_int, [_vp, _vp, _vp, _u32, _f, _f, _vp, _vp]),
    ("aurora_trace_closest_device", C.c_int, [_vp, _vp, _vp, _u32, _f, _f, _vp, _vp, _vp]),
    ("aurora_trace_closest_device_noalloc", C.c_int, [_vp, _vp, _vp, _u32, _f, _f, _vp, _vp, _vp]),
    ("aurora_trace_brute_device", C.c_int, [_vp, _vp, _vp, _u32, _f, _f, _vp, _vp]),
    ("aurora_trace_closest_device_stats", C.c_int,
     [_vp, _vp, _vp, _u32, _f, _f, _vp, _vp, C.c_int, C.POINTER(TraceStats)]),
    ("aurora_render", C.c_int,
     [_vp, C.POINTER(CameraDesc), C.POINTER(RenderDesc), _vp, _vp, _vp, C.POINTER(RenderStats)]),
    ("aurora_render_device", C.c_int,
     [_vp, C.POINTER(CameraDesc), C.POINTER(RenderDesc), _vp, _vp, _vp, C.POINTER(RenderStats)]),
    ("aurora_camera_rays", C.c_int, [_vp, C.POINTER(CameraDesc), _u32, _vp, _vp, _u32, _vp, _vp]),
    ("aurora_camera_rays_device", C.c_int, [_vp, C.POINTER(CameraDesc), _u32, _vp, _vp, _u32, _vp, _vp]),
    ("aurora_paste", C.c_int, [_vp, _vp, _u32, _u32, _vp, _u32, _u32, _vp]),
    ("aurora_paste_device", C.c_int, [_vp, _vp, _u32, _u32, _vp, _u32, _u32, _vp]),
    ("aurora_patch_render", C.c_int, [_vp, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _vp]),
    ("aurora_patch_render_device", C.c_int, [_vp, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _vp]),
    ("aurora_resolve_rgb8", C.c_int, [_vp, _u32, _u32, _u32, _u32, _vp]),
    ("aurora_resolve_rgb8_device", C.c_int, [_vp, _vp, _u32, _u32, _u32, _u32, _vp]),
    ("aurora_render_batch_device", C.c_int, [_vp, C.POINTER(BatchItem), _u32, _u32, C.POINTER(RenderStats)]),
    ("aurora_denoiser_create", C.c_int, [_vp, _u32, C.POINTER(_vp)]),
    ("aurora_denoiser_destroy", None, [_vp]),
    ("aurora_denoiser_param_count", C.c_uint64, [_vp]),
    ("aurora_denoiser_get_params", C.c_int, [_vp, _vp, C.c_uint64]),
    ("aurora_denoiser_set_params", C.c_int, [_vp, _vp, C.c_uint64]),
    ("aurora_denoise", C.c_int, [_vp, _vp, _vp, _vp, _u32, _u32, _u32, _vp]),
    ("aurora_denoiser_set_loss", C.c_int, [_vp, C.c_int]),
    ("aurora_denoiser_loss_grad", C.c_int,
     [_vp, _vp, _vp, _vp, _vp, _u32, _u32, _u32, _u32, C.POINTER(C.c_float), _vp]),
    ("aurora_denoiser_train_step", C.c_int,
     [_vp, _vp, _vp, _vp, _vp, _u32, _u32, _u32, _u32, _f, C.POINTER(C.c_float)]),
    ("aurora_selftest_mfma", C.c_int, [_vp, _vp, _vp, _vp]),
    ("aurora_selftest_rcp", C.c_int, [_vp, C.c_uint32, C.c_uint32, _vp, _vp]),
]

_lib = None


def load_library(path=None):
    """Load libaurora.so (built in-tree by __graft_entry__.build()).

    Raises RuntimeError when the library is missing: there is no CPU
    fallback for the render path."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    # PyTorch ships its own HIP runtime; importing it first makes that the
    # one runtime of the process, which libaurora then shares (loading
    # libaurora first and torch second leaves torch without a device)
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(p):
        raise RuntimeError(f"libaurora not built: {p} is missing (run __graft_entry__.build())")
    lib = C.CDLL(p)
    for name, res, args in EXPORTS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.aurora_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{p} has ABI version {lib.aurora_abi_version()}, this binding expects {ABI_VERSION} "
                           "(rebuild with __graft_entry__.build())")
    if path is None:
        _lib = lib
    return lib
