"""ctypes mirror of include/rt_scene.h and include/rt_abi.h, and the loader of
librt_hip.so.  Loading fails loudly when the library has not been built."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librt_hip.so")

c_float3 = ctypes.c_float * 3


class RtMaterial(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("kd", c_float3), ("ks", c_float3), ("exponent", ctypes.c_float),
                ("texture", ctypes.c_int32)]


class RtLight(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("p", c_float3), ("u", c_float3), ("v", c_float3),
                ("intensity", c_float3)]


class RtSceneDesc(ctypes.Structure):
    _fields_ = [
        ("n_vertices", ctypes.c_int32),
        ("positions", ctypes.c_void_p),
        ("normals", ctypes.c_void_p),
        ("n_triangles", ctypes.c_int32),
        ("tri_vertices", ctypes.c_void_p),
        ("tri_material", ctypes.c_void_p),
        ("tri_prim_id", ctypes.c_void_p),
        ("n_spheres", ctypes.c_int32),
        ("spheres", ctypes.c_void_p),
        ("sphere_material", ctypes.c_void_p),
        ("sphere_prim_id", ctypes.c_void_p),
        ("n_materials", ctypes.c_int32),
        ("materials", ctypes.c_void_p),
        ("n_lights", ctypes.c_int32),
        ("lights", ctypes.c_void_p),
        ("env_width", ctypes.c_int32),
        ("env_height", ctypes.c_int32),
        ("env_rgb", ctypes.c_void_p),
        ("env_kind", ctypes.c_int32),
        ("n_objects", ctypes.c_int32),
        ("csg_nodes", ctypes.c_void_p),
        ("object_first", ctypes.c_void_p),
        ("object_prim_id", ctypes.c_void_p),
        ("n_textures", ctypes.c_int32),
        ("textures", ctypes.c_void_p),
        ("tri_texcoords", ctypes.c_void_p),
        ("object_xform", ctypes.c_void_p),
        ("bvh_build", ctypes.c_int32),
    ]


class RtTexture(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("rgb", ctypes.c_void_p)]


class RtCsgNode(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int32), ("material", ctypes.c_int32), ("p", ctypes.c_float * 6)]


CSG_SPHERE, CSG_BOX, CSG_PLANE, CSG_ISECT, CSG_UNION, CSG_DIFF, CSG_XOR = range(7)
CSG_MAX_LEAVES = 16


class RtCamera(ctypes.Structure):
    _fields_ = [("eye", c_float3), ("look_at", c_float3), ("up", c_float3), ("fov", ctypes.c_float),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32)]


class RtView(ctypes.Structure):
    _fields_ = [("camera", RtCamera), ("spp_x", ctypes.c_int32), ("spp_y", ctypes.c_int32)]


class RtParams(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("max_depth", ctypes.c_int32), ("rr_start", ctypes.c_int32),
                ("rr_prob", ctypes.c_float), ("seed", ctypes.c_uint32), ("guide_train", ctypes.c_int32),
                ("guide_lr", ctypes.c_float), ("guide_batch", ctypes.c_int32)]


class RtStats(ctypes.Structure):
    _fields_ = [
        ("camera_rays", ctypes.c_int64),
        ("extension_rays", ctypes.c_int64),
        ("shadow_rays", ctypes.c_int64),
        ("render_ms", ctypes.c_double),
        ("extend_ms", ctypes.c_double),
        ("shadow_ms", ctypes.c_double),
        ("extend_launches", ctypes.c_int64),
        ("bvh_nodes", ctypes.c_int64),
        ("bvh_prims", ctypes.c_int64),
        ("bvh_depth", ctypes.c_int32),
        ("paths_per_pass", ctypes.c_int32),
        ("passes", ctypes.c_int32),
        ("bvh_stack_need", ctypes.c_int32),
        ("node_visits", ctypes.c_int64),
        ("prim_tests", ctypes.c_int64),
        ("guide_infer_ms", ctypes.c_double),
        ("guide_train_ms", ctypes.c_double),
        ("guide_infer_samples", ctypes.c_int64),
        ("guide_train_samples", ctypes.c_int64),
        ("guide_train_steps", ctypes.c_int64),
        ("tonemap_ms", ctypes.c_double),
        ("shadow_node_visits", ctypes.c_int64),
        ("shadow_prim_tests", ctypes.c_int64),
        ("node_visits_lds", ctypes.c_int64),
        ("shadow_node_visits_lds", ctypes.c_int64),
        ("finish_ms", ctypes.c_double),
        ("finish_rays", ctypes.c_int64),
        ("finish_shadow_rays", ctypes.c_int64),
        ("finish_bounce", ctypes.c_int64),
    ]


class RtLaunchConfig(ctypes.Structure):
    _fields_ = [("path_slots", ctypes.c_int32), ("block", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32),
                ("use_graph", ctypes.c_int32),
                ("count_traversal", ctypes.c_int32), ("guide_epochs", ctypes.c_int32),
                ("finish_depth", ctypes.c_int32), ("finish_paths", ctypes.c_int32)]


class RtMltParams(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("max_depth", ctypes.c_int32), ("rr_start", ctypes.c_int32),
                ("termination_prob", ctypes.c_float), ("large_step_prob", ctypes.c_float),
                ("bootstrap", ctypes.c_int32), ("mutations", ctypes.c_int32), ("chains", ctypes.c_int32),
                ("erpt_spp_x", ctypes.c_int32), ("erpt_spp_y", ctypes.c_int32),
                ("erpt_ed_termination", ctypes.c_float), ("seed", ctypes.c_uint32),
                ("flt_passes", ctypes.c_int32), ("flt_flags", ctypes.c_int32)]


class RtMltResult(ctypes.Structure):
    _fields_ = [("b", ctypes.c_double), ("ed", ctypes.c_double), ("chains", ctypes.c_int64),
                ("mutations", ctypes.c_int64), ("accepted", ctypes.c_int64), ("seed_paths", ctypes.c_int64),
                ("rays", ctypes.c_int64), ("render_ms", ctypes.c_double),
                ("chain_seeds", ctypes.c_int64)]


MLT_METROPOLIS, MLT_ERPT, MLT_FILTERED = 0, 1, 2
FLT_NO_FILTER, FLT_NO_CROSSFILTERING, FLT_NO_VARIANCE_FILTER, FLT_SIMPLE_VARIANCE_FILTER = 1, 2, 4, 8
MAT_DIFFUSE, MAT_MIRROR, MAT_PHONG, MAT_DARK, MAT_BLINN, MAT_REFRACTIVE, MAT_TEXTURE = 0, 1, 2, 3, 4, 5, 6
LIGHT_AREA, LIGHT_POINT, LIGHT_ENV, LIGHT_CONE = 0, 1, 2, 3
MODE_COMPAT, MODE_UNBIASED, MODE_GUIDED = 0, 1, 2
BVH_MEDIAN, BVH_SAH = 0, 1

STATUS = {0: "ok", -1: "invalid argument", -2: "no scene uploaded", -3: "HIP runtime error",
          -4: "BVH build failed", -5: "cancelled", -6: "injected fault", -7: "no guiding network",
          -8: "scene parse error", -9: "unsupported scene construct", -10: "file error"}
ERR_INVALID, ERR_PARSE, ERR_UNSUPPORTED, ERR_IO = -1, -8, -9, -10

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32

# name -> (restype, argtypes); every symbol declared in include/rt_abi.h
PROTOTYPES = {
    "rt_error_string": (ctypes.c_char_p, [ctypes.c_int]),
    "rt_version": (ctypes.c_int, []),
    "rt_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "rt_destroy": (ctypes.c_int, [_P]),
    "rt_last_error": (ctypes.c_char_p, [_P]),
    "rt_scene_upload": (ctypes.c_int, [_P, ctypes.POINTER(RtSceneDesc)]),
    "rt_trace_closest": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P]),
    "rt_trace_closest_device": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P]),
    "rt_trace_any": (ctypes.c_int, [_P, _P, _P, _I64, _P]),
    "rt_trace_any_device": (ctypes.c_int, [_P, _P, _P, _I64, _P]),
    "rt_trace_count_device": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P]),
    "rt_trace_any_count_device": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P]),
    "rt_bvh_export": (ctypes.c_int, [ctypes.POINTER(RtSceneDesc), _P, ctypes.POINTER(ctypes.c_int64), _P,
                                     ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32)]),
    "rt_trace_closest_device_noalloc": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P]),
    "rt_render": (ctypes.c_int, [_P, ctypes.POINTER(RtView), _I32, ctypes.POINTER(RtParams), _P]),
    "rt_render_device": (ctypes.c_int, [_P, ctypes.POINTER(RtView), _I32, ctypes.POINTER(RtParams), _P]),
    "rt_integrate_rays": (ctypes.c_int, [_P, _P, _P, _I64, ctypes.POINTER(RtParams), ctypes.c_uint32, _P]),
    "rt_integrate_rays_device": (ctypes.c_int, [_P, _P, _P, _I64, ctypes.POINTER(RtParams), ctypes.c_uint32, _P]),
    "rt_set_launch_config": (ctypes.c_int, [_P, ctypes.POINTER(RtLaunchConfig)]),
    "rt_get_launch_config": (ctypes.c_int, [_P, ctypes.POINTER(RtLaunchConfig)]),
    "rt_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(RtStats)]),
    "rt_request_stop": (ctypes.c_int, [_P, _I32]),
    "rt_set_fault": (ctypes.c_int, [_P, _I32]),
    "rt_guide_param_count": (ctypes.c_int, []),
    "rt_guide_init": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "rt_guide_set_params": (ctypes.c_int, [_P, _P, _I32]),
    "rt_guide_get_params": (ctypes.c_int, [_P, _P, _I32]),
    "rt_guide_set_bounds": (ctypes.c_int, [_P, _P, _P]),
    "rt_guide_forward": (ctypes.c_int, [_P, _P, _I32, _P]),
    "rt_guide_head": (ctypes.c_int, [_P, _P, _I32, _P]),
    "rt_guide_gradient": (ctypes.c_int, [_P, _P, _I32, _P, _P]),
    "rt_guide_train": (ctypes.c_int, [_P, _P, _I32, _I32, ctypes.c_float]),
    # include/rt_scene_file.h
    "rt_scene_file_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    "rt_scene_file_parse": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    "rt_scene_file_last_error": (ctypes.c_char_p, []),
    "rt_scene_file_messages": (ctypes.c_char_p, [_P]),
    "rt_scene_file_scene": (ctypes.c_int, [_P, ctypes.POINTER(RtSceneDesc), ctypes.POINTER(RtCamera)]),
    "rt_scene_file_dump_bvh": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "rt_scene_file_dump_bvh_text": (_I64, [_P, _P, _I64]),
    "rt_scene_file_free": (None, [_P]),
    # include/rt_film.h
    "rt_film_tonemap": (ctypes.c_int, [_P, _P, _I32, _I32, ctypes.c_float, _P]),
    "rt_film_tonemap_device": (ctypes.c_int, [_P, _P, _I32, _I32, ctypes.c_float, _P]),
    "rt_film_write_exr": (ctypes.c_int, [ctypes.c_char_p, _P, _I32, _I32]),
    "rt_film_write_image": (ctypes.c_int, [ctypes.c_char_p, _P, _I32, _I32]),
    "rt_film_write_mat": (ctypes.c_int, [ctypes.c_char_p, _P, _I32, _I32]),
    "rt_unique_filename": (_I64, [ctypes.c_char_p, _P, _I64]),
    "rt_film_last_error": (ctypes.c_char_p, []),
    "rt_image_read_hdr": (ctypes.c_int, [ctypes.c_char_p, _P, _P, _P, _I64]),
    "rt_image_read": (ctypes.c_int, [ctypes.c_char_p, _P, _P, _P, _I64]),
    # include/rt_mlt.h
    "rt_integrate_samples": (ctypes.c_int, [_P, ctypes.POINTER(RtCamera), _I32, _I32, ctypes.c_float, _P, _I64, _P]),
    "rt_render_mlt": (ctypes.c_int, [_P, ctypes.POINTER(RtCamera), ctypes.POINTER(RtMltParams), _P,
                                     ctypes.POINTER(RtMltResult)]),
    "rt_render_mlt_device": (ctypes.c_int, [_P, ctypes.POINTER(RtCamera), ctypes.POINTER(RtMltParams), _P,
                                            ctypes.POINTER(RtMltResult)]),
    "rt_symmetric_filter": (ctypes.c_int, [_P, _P, _P, _P, _I32, _I32, _I32, _I32, _P]),
}

_lib = None


class RtError(RuntimeError):
    def __init__(self, status, msg=""):
        super().__init__(f"librt_hip: {STATUS.get(status, status)} ({status}){': ' + msg if msg else ''}")
        self.status = status


def load():
    """Load librt_hip.so (built by renderer_amd.build); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m renderer_amd.build` "
                           "(the GPU path has no fallback)")
    try:
        # torch ships its own HIP runtime under an unversioned file name; loaded
        # after this library's (/opt/rocm) it would be a second runtime that sees
        # no GPU.  Loading torch first makes this library bind to torch's copy,
        # so both share one runtime.
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
