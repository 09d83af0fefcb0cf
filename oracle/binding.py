"""ctypes binding of the CPU checker (oracle/build/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker - never by the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from kiraray_amd._abi import krr_camera, krr_hit, krr_material, krr_medium, krr_scene_desc

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")


class oracle_params(C.Structure):
    _fields_ = [("nee", C.c_int32), ("enable_medium", C.c_int32), ("max_depth", C.c_int32), ("rr", C.c_float),
                ("rr_min_depth", C.c_int32), ("enable_clamp", C.c_int32), ("clamp_max", C.c_float),
                ("spp", C.c_int32), ("legacy_spp_accumulate", C.c_int32), ("precision", C.c_int32),
                ("n_threads", C.c_int32), ("row_begin", C.c_int32), ("row_end", C.c_int32),
                ("row_step", C.c_int32), ("pixel_step", C.c_int32), ("view_id", C.c_int32)]


class oracle_stats(C.Structure):
    _fields_ = [("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("paths", C.c_uint64),
                ("seconds", C.c_double), ("threads", C.c_int32)]


F3 = C.c_float * 3
F2 = C.c_float * 2
_PROTO = {
    "oracle_default_params": (None, [C.POINTER(oracle_params)]),
    "oracle_render": (C.c_int, [C.POINTER(krr_scene_desc), C.POINTER(krr_camera), C.POINTER(oracle_params), C.c_uint64,
                                C.c_int32, C.c_int32, C.c_void_p, C.POINTER(oracle_stats)]),
    "oracle_render_megakernel": (C.c_int, [C.POINTER(krr_scene_desc), C.POINTER(krr_camera), C.POINTER(oracle_params),
                                           C.c_uint64, C.c_int32, C.c_int32, C.c_void_p, C.POINTER(oracle_stats)]),
    "oracle_pcg32_sequence": (None, [C.c_uint64, C.c_uint64, C.c_int32, C.c_void_p]),
    "oracle_pcg32_advance": (None, [C.c_uint64, C.c_uint64, C.c_int64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "oracle_pixel_samples": (None, [C.c_int32, C.c_int32, C.c_uint32, C.c_int64, C.c_int32, C.c_void_p]),
    "oracle_interleave_32bit": (C.c_uint32, [C.c_uint32, C.c_uint32]),
    "oracle_murmur64a": (C.c_uint64, [C.c_void_p, C.c_uint64, C.c_uint64]),
    "oracle_trace_closest": (C.c_int, [C.POINTER(krr_scene_desc), C.c_void_p, C.c_int64, C.c_void_p, C.c_int32]),
    "oracle_trace_any": (C.c_int, [C.POINTER(krr_scene_desc), C.c_void_p, C.c_int64, C.c_void_p, C.c_int32]),
    "oracle_camera_ray": (None, [C.POINTER(krr_camera), C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                 C.c_void_p, C.c_void_p]),
    "oracle_bsdf_eval": (None, [C.POINTER(krr_material), C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    "oracle_bsdf_sample": (C.c_int32, [C.POINTER(krr_material), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.POINTER(C.c_float)]),
    "oracle_bsdf_type": (C.c_int32, [C.POINTER(krr_material)]),
    "oracle_fresnel_dielectric": (C.c_float, [C.c_float, C.c_float]),
    "oracle_fresnel_complex": (None, [C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]),
    "oracle_ggx_D": (C.c_float, [C.c_float, C.c_float, C.c_void_p]),
    "oracle_ggx_G1": (C.c_float, [C.c_float, C.c_float, C.c_void_p]),
    "oracle_ggx_pdf": (C.c_float, [C.c_float, C.c_float, C.c_void_p, C.c_void_p]),
    "oracle_ggx_sample": (None, [C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]),
    "oracle_hg_p": (C.c_float, [C.c_float, C.c_void_p, C.c_void_p]),
    "oracle_hg_sample": (None, [C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    "oracle_sample_discrete3": (C.c_int32, [C.c_void_p, C.c_float, C.POINTER(C.c_float)]),
    "oracle_majorant_grid": (C.c_int, [C.POINTER(krr_medium), C.c_void_p]),
    "oracle_density": (C.c_float, [C.POINTER(krr_medium), C.c_void_p]),
    "oracle_dm_sin": (C.c_float, [C.c_float]),
    "oracle_dm_cos": (C.c_float, [C.c_float]),
    "oracle_dm_log": (C.c_float, [C.c_float]),
    "oracle_dm_exp": (C.c_float, [C.c_float]),
    "oracle_dm_tan": (C.c_float, [C.c_float]),
    "oracle_dm_atan2": (C.c_float, [C.c_float, C.c_float]),
    "oracle_error_metric": (C.c_double, [C.c_int32, C.c_void_p, C.c_void_p, C.c_int64]),
    "oracle_majorant_segments": (C.c_int32, [C.POINTER(krr_medium), C.c_void_p, C.c_void_p, C.c_float, C.c_int32,
                                             C.c_int32, C.c_void_p]),
    "oracle_transmittance": (C.c_int, [C.POINTER(krr_scene_desc), C.c_void_p, C.c_void_p, C.c_float, C.c_int32,
                                       C.c_int32, C.c_float, C.c_float, C.c_uint64, C.c_int64, C.c_int32, C.c_void_p,
                                       C.c_void_p]),
}

_lib = None


def build():
    """Compile liboracle.so with its Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def _bind(path):
    lib = C.CDLL(path)
    for n, (res, args) in _PROTO.items():
        f = getattr(lib, n)
        f.restype = res
        f.argtypes = args
    return lib


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _bind(LIB_PATH)
    return _lib


def load_planted(k):
    """The checker built with planted mistake k (ORACLE_PLANT in core.h):
    a separate library, for tests that show a pin fails on that mistake."""
    path = os.path.join(HERE, "build", f"liboracle_plant{int(k)}.so")
    subprocess.run(["make", "-s", "-C", HERE, "plant", f"PLANT={int(k)}"], check=True)
    return _bind(path)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def params(lib=None, **kw):
    p = oracle_params()
    (lib or load()).oracle_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def render(desc, cam, width, height, frame_index=1, film=None, lib=None, **kw):
    """One frame (beginFrame + render) -> (film[h, w, 4] float32, stats dict).
    lib: another build of the checker (load_planted), default the checker."""
    lib = lib or load()
    if film is None:
        film = np.zeros((height, width, 4), np.float32)
    st = oracle_stats()
    p = params(lib=lib, **kw)
    rc = lib.oracle_render(C.byref(desc), C.byref(cam), C.byref(p), frame_index, width, height, _ptr(film), C.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle_render failed: {rc}")
    return film, {"closest_rays": st.closest_rays, "shadow_rays": st.shadow_rays, "paths": st.paths,
                  "seconds": st.seconds, "threads": st.threads}


def render_megakernel(desc, cam, width, height, frame_index=1, film=None, **kw):
    """One frame of the megakernel path tracer -> (film[h, w, 4] float32, stats dict)."""
    lib = load()
    if film is None:
        film = np.zeros((height, width, 4), np.float32)
    st = oracle_stats()
    p = params(**kw)
    rc = lib.oracle_render_megakernel(C.byref(desc), C.byref(cam), C.byref(p), frame_index, width, height, _ptr(film),
                                      C.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle_render_megakernel failed: {rc}")
    return film, {"closest_rays": st.closest_rays, "shadow_rays": st.shadow_rays, "paths": st.paths,
                  "seconds": st.seconds, "threads": st.threads}


def pcg32_sequence(initstate, initseq, n):
    out = np.zeros(n, np.uint32)
    load().oracle_pcg32_sequence(initstate, initseq, n, _ptr(out))
    return out


def pcg32_advance(initstate, initseq, delta):
    s, i = C.c_uint64(), C.c_uint64()
    load().oracle_pcg32_advance(initstate, initseq, delta, C.byref(s), C.byref(i))
    return s.value, i.value


def pixel_samples(x, y, sample_index, adv, n):
    out = np.zeros(n, np.float32)
    load().oracle_pixel_samples(x, y, sample_index, adv, n, _ptr(out))
    return out


def trace_closest(desc, rays, use_bvh=False):
    rays = np.ascontiguousarray(rays, np.float32)
    hits = np.zeros(len(rays), dtype=[("t", "<f4"), ("prim", "<i4"), ("b1", "<f4"), ("b2", "<f4")])
    rc = load().oracle_trace_closest(C.byref(desc), _ptr(rays), len(rays), _ptr(hits), 1 if use_bvh else 0)
    assert rc == 0
    return hits


def trace_any(desc, rays, use_bvh=False):
    rays = np.ascontiguousarray(rays, np.float32)
    vis = np.zeros(len(rays), np.int32)
    rc = load().oracle_trace_any(C.byref(desc), _ptr(rays), len(rays), _ptr(vis), 1 if use_bvh else 0)
    assert rc == 0
    return vis


def camera_ray(cam, x, y, w, h, sample):
    s = np.asarray(sample, np.float32)
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    load().oracle_camera_ray(C.byref(cam), x, y, w, h, _ptr(s), _ptr(o), _ptr(d))
    return o, d


def material(mtype, diffuse=(1, 1, 1), roughness=1.0, ior=1.5, spec_trans=0.0, anisotropic=0.0):
    m = krr_material()
    m.type = mtype
    for i in range(3):
        m.diffuse[i] = diffuse[i]
    m.specular[3] = 1.0 - roughness
    m.ior = ior
    m.specular_transmission = spec_trans
    m.anisotropic = anisotropic
    return m


def bsdf_eval(m, wo, wi):
    wo = np.asarray(wo, np.float32)
    wi = np.asarray(wi, np.float32)
    f = np.zeros(3, np.float32)
    pdf = C.c_float()
    load().oracle_bsdf_eval(C.byref(m), _ptr(wo), _ptr(wi), _ptr(f), C.byref(pdf))
    return f, pdf.value


def bsdf_sample(m, wo, u):
    wo = np.asarray(wo, np.float32)
    u = np.asarray(list(u) + [0.0] * (3 - len(u)), np.float32)
    f = np.zeros(3, np.float32)
    wi = np.zeros(3, np.float32)
    pdf = C.c_float()
    flags = load().oracle_bsdf_sample(C.byref(m), _ptr(wo), _ptr(u), _ptr(f), _ptr(wi), C.byref(pdf))
    return f, wi, pdf.value, flags


def murmur64a(data, seed=0):
    buf = np.frombuffer(bytes(data), np.uint8).copy()
    return load().oracle_murmur64a(_ptr(buf), len(buf), seed)


def error_metric(metric, img, ref):
    img = np.ascontiguousarray(img, np.float32)
    ref = np.ascontiguousarray(ref, np.float32)
    return load().oracle_error_metric(metric, _ptr(img), _ptr(ref), img.size // 4)


def majorant_segments(medium, o, d, tmax, precision=64, cap=4096, lib=None):
    """sampleRay + MajorantIterator walk of medium (a krr_medium) along
    o + t d, t in [0, tmax] -> array (n, 5): tMin, tMax, sigma_maj rgb."""
    seg = np.zeros((cap, 5), np.float64)
    o = np.asarray(o, np.float32)
    d = np.asarray(d, np.float32)
    n = (lib or load()).oracle_majorant_segments(C.byref(medium), _ptr(o), _ptr(d), float(tmax), precision, cap, _ptr(seg))
    assert 0 <= n <= cap, n
    return seg[:n]


def transmittance(desc, o, d, tmax, medium, channel, n, seed=1, sr_pu=0.0, sr_pl=1.0, precision=64, lib=None):
    """Mean and standard error (per channel) of traceTransmittance's
    contribution for the shadow ray {o, d, tmax, medium}, Ld = 1, over n runs."""
    mean = np.zeros(3, np.float64)
    sq = np.zeros(3, np.float64)
    o = np.asarray(o, np.float32)
    d = np.asarray(d, np.float32)
    rc = (lib or load()).oracle_transmittance(C.byref(desc), _ptr(o), _ptr(d), float(tmax), medium, channel, sr_pu, sr_pl,
                                              seed, n, precision, _ptr(mean), _ptr(sq))
    assert rc == 0, rc
    se = np.sqrt(np.maximum(sq - mean ** 2, 0.0) / n)
    return mean, se
