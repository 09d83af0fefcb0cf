"""ctypes declarations of include/krr_wavefront.h.

Every structure mirrors the C header field for field; every function gets
explicit argtypes/restype so that a mismatch fails loudly instead of passing
garbage across the boundary.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libkrr_hip.so")
NRC_PARAMS = 5 * 64 * 64 + 64 * 3  # KRR_NRC_PARAMS

KRR_OK = 0
KRR_ERR_INVALID_ARG = -1
KRR_ERR_BAD_CONFIG = -2
KRR_ERR_NO_SCENE = -3
KRR_ERR_HIP = -4
KRR_ERR_OUT_OF_MEMORY = -5
KRR_ERR_CAPACITY = -6
KRR_ERR_UNSUPPORTED = -7
KRR_ERR_NO_DEVICE = -8

KRR_MAT_NULL, KRR_MAT_DIFFUSE, KRR_MAT_DIELECTRIC, KRR_MAT_CONDUCTOR = 0, 1, 2, 3
KRR_MEDIUM_HOMOGENEOUS, KRR_MEDIUM_GRID = 0, 1
KRR_METRIC_MSE, KRR_METRIC_MAPE, KRR_METRIC_SMAPE, KRR_METRIC_RELMSE = 0, 1, 2, 3


class krr_material(C.Structure):
    _fields_ = [("type", C.c_int32), ("diffuse", C.c_float * 3), ("specular", C.c_float * 4),
                ("ior", C.c_float), ("specular_transmission", C.c_float), ("anisotropic", C.c_float),
                ("diffuse_texture", C.c_int32), ("specular_texture", C.c_int32),
                ("transmission_texture", C.c_int32), ("transmission", C.c_float * 3),
                ("emissive_texture", C.c_int32), ("normal_texture", C.c_int32)]


class krr_texture(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("rgba", C.POINTER(C.c_float))]


class krr_mesh(C.Structure):
    _fields_ = [("first_triangle", C.c_int32), ("n_triangles", C.c_int32), ("material", C.c_int32),
                ("medium_inside", C.c_int32), ("medium_outside", C.c_int32), ("has_normals", C.c_int32),
                ("emission", C.c_float * 3), ("has_texcoords", C.c_int32), ("has_tangents", C.c_int32)]


class krr_medium(C.Structure):
    _fields_ = [("type", C.c_int32), ("sigma_t", C.c_float * 3), ("albedo", C.c_float * 3),
                ("Le", C.c_float * 3), ("g", C.c_float), ("scale", C.c_float), ("res", C.c_int32 * 3),
                ("density", C.POINTER(C.c_float)), ("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3)]


class krr_infinite_light(C.Structure):
    _fields_ = [("tint", C.c_float * 3), ("scale", C.c_float), ("rotation", C.c_float * 9),
                ("image_width", C.c_int32), ("image_height", C.c_int32), ("image_rgba", C.POINTER(C.c_float))]


class krr_scene_desc(C.Structure):
    _fields_ = [("n_vertices", C.c_int64), ("positions", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("n_triangles", C.c_int64), ("indices", C.POINTER(C.c_int32)),
                ("n_meshes", C.c_int32), ("meshes", C.POINTER(krr_mesh)),
                ("n_materials", C.c_int32), ("materials", C.POINTER(krr_material)),
                ("n_media", C.c_int32), ("media", C.POINTER(krr_medium)),
                ("n_infinite_lights", C.c_int32), ("infinite_lights", C.POINTER(krr_infinite_light)),
                ("texcoords", C.POINTER(C.c_float)), ("n_textures", C.c_int32), ("textures", C.POINTER(krr_texture)),
                ("tangents", C.POINTER(C.c_float))]


class krr_camera(C.Structure):
    _fields_ = [("film_size", C.c_float * 2), ("focal_length", C.c_float), ("focal_distance", C.c_float),
                ("lens_radius", C.c_float), ("aspect_ratio", C.c_float), ("shutter_open", C.c_float),
                ("shutter_time", C.c_float), ("transform", C.c_float * 12), ("medium", C.c_int32)]


class krr_view(C.Structure):
    _fields_ = [("camera", krr_camera), ("width", C.c_int32), ("height", C.c_int32), ("film_dev", C.c_void_p)]


class krr_stats(C.Structure):
    _fields_ = [("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("paths", C.c_uint64),
                ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64), ("shadow_node_visits", C.c_uint64),
                ("shadow_tri_tests", C.c_uint64), ("closest_launches", C.c_uint64),
                ("closest_trace_ms", C.c_double), ("shade_ms", C.c_double), ("shadow_trace_ms", C.c_double),
                ("render_ms", C.c_double), ("megakernel_ms", C.c_double),
                ("megakernel_closest_rays", C.c_uint64), ("megakernel_shadow_rays", C.c_uint64),
                ("megakernel_node_visits", C.c_uint64),
                ("megakernel_tri_tests", C.c_uint64),
                ("nrc_ms", C.c_double), ("nrc_queries", C.c_uint64), ("nrc_train_vertices", C.c_uint64),
                ("nrc_steps", C.c_uint64), ("nrc_loss", C.c_double), ("bvh_build_ms", C.c_double)]


class krr_hit(C.Structure):
    _fields_ = [("t", C.c_float), ("prim", C.c_int32), ("b1", C.c_float), ("b2", C.c_float)]


# name -> (restype, argtypes)
PROTOTYPES = {
    "krr_last_error": (C.c_char_p, []),
    "krr_abi_version": (C.c_int, []),
    "krr_device_count": (C.c_int, []),
    "krr_scene_create": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "krr_scene_get_desc": (C.c_int, [C.c_void_p, C.POINTER(krr_scene_desc)]),
    "krr_scene_get_camera": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(krr_camera)]),
    "krr_scene_view_count": (C.c_int, [C.c_void_p]),
    "krr_scene_view_size": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "krr_scene_destroy": (C.c_int, [C.c_void_p]),
    "krr_accel_build": (C.c_int, [C.POINTER(krr_scene_desc), C.c_char_p, C.POINTER(C.c_void_p)]),
    "krr_accel_trace_closest": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "krr_accel_trace_any": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "krr_accel_trace_closest_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "krr_accel_noalloc_scratch_bytes": (C.c_size_t, []),
    "krr_accel_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "krr_accel_destroy": (C.c_int, [C.c_void_p]),
    "krr_wf_create": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "krr_wf_set_scene": (C.c_int, [C.c_void_p, C.POINTER(krr_scene_desc)]),
    "krr_wf_resize": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "krr_wf_begin_frame": (C.c_int, [C.c_void_p, C.POINTER(krr_camera), C.c_uint64]),
    "krr_wf_render": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "krr_wf_end_frame": (C.c_int, [C.c_void_p]),
    "krr_wf_stats": (C.c_int, [C.c_void_p, C.POINTER(krr_stats)]),
    "krr_wf_get_params": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32]),
    "krr_wf_destroy": (C.c_int, [C.c_void_p]),
    "krr_mk_create": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "krr_mk_set_scene": (C.c_int, [C.c_void_p, C.POINTER(krr_scene_desc)]),
    "krr_mk_resize": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "krr_mk_begin_frame": (C.c_int, [C.c_void_p, C.POINTER(krr_camera), C.c_uint64]),
    "krr_mk_render": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "krr_mk_stats": (C.c_int, [C.c_void_p, C.POINTER(krr_stats)]),
    "krr_mk_get_params": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32]),
    "krr_mk_destroy": (C.c_int, [C.c_void_p]),
    "krr_wf_set_views": (C.c_int, [C.c_void_p, C.POINTER(krr_view), C.c_int32]),
    "krr_wf_render_views": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "krr_wf_autotune": (C.c_int, [C.c_void_p, C.c_int64, C.c_char_p, C.c_int32, C.c_void_p]),
    "krr_error_metric": (C.c_int, [C.c_int32, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_double), C.c_void_p]),
    "krr_accumulate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_int32, C.c_int32, C.c_void_p]),
    # neural radiance cache (include/krr_nrc.h)
    "krr_nrc_create": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "krr_nrc_destroy": (C.c_int, [C.c_void_p]),
    "krr_nrc_get_weights": (C.c_int, [C.c_void_p, C.c_void_p]),
    "krr_nrc_set_weights": (C.c_int, [C.c_void_p, C.c_void_p]),
    "krr_nrc_encode": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "krr_nrc_infer": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "krr_nrc_gradient": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(C.c_float),
                                   C.c_void_p]),
    "krr_nrc_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_float), C.c_void_p]),
    "krr_nrc_steps": (C.c_int64, [C.c_void_p]),
}

_lib = None


def load_library(path=None):
    """Load libkrr_hip.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"kiraray_amd: HIP library {p} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(p)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
