"""kiraray_amd - MI355X (gfx950) wavefront path tracer.

Python mirror of KiRaRay's WavefrontPathTracer render pass
(src/render/wavefront/integrator.h:24-104) over the C ABI of
include/krr_wavefront.h.  The pass keeps the reference's hook order
(setScene -> resize -> per frame beginFrame / render / endFrame), its JSON
params (nee, enable_medium, max_depth, rr, enable_clamp, clamp_max plus spp
and the additions listed in the header) and its film layout: RGBA32F,
width x height float4, pixel row y stored at row height-1-y.

Device memory and streams can come from PyTorch (``film.data_ptr()``,
``torch.cuda.current_stream().cuda_stream``); nothing here falls back to a
CPU path - every compute call goes to the HIP library and raises if it is
missing.
"""
import ctypes as C
import json as _json

from . import _abi
from ._abi import (krr_camera, krr_hit, krr_material, krr_medium, krr_mesh, krr_scene_desc, krr_stats, krr_view,
                   load_library)

__all__ = ["KrrError", "Scene", "Accel", "WavefrontPathTracer", "MegakernelPathTracer", "RenderApp", "load_library", "error_metric", "accumulate",
           "krr_camera", "krr_hit", "krr_material", "krr_medium", "krr_mesh", "krr_scene_desc", "krr_stats", "krr_view"]


class KrrError(RuntimeError):
    """A negative krr_status from the library, with krr_last_error()."""

    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


def _check(rc, lib=None):
    if rc < 0:
        lib = lib or load_library()
        raise KrrError(rc, lib.krr_last_error().decode(errors="replace"))
    return rc


def _json_arg(cfg):
    if cfg is None:
        return b""
    if isinstance(cfg, (bytes, str)):
        return cfg.encode() if isinstance(cfg, str) else cfg
    return _json.dumps(cfg).encode()


class Scene:
    """Procedural scene (SceneImporter counterpart), host memory only.

    The arrays that `desc` points to belong to the scene and are freed with
    it: keep the Scene alive while the descriptor is read."""

    def __init__(self, config):
        self._lib = load_library()
        h = C.c_void_p()
        _check(self._lib.krr_scene_create(_json_arg(config), C.byref(h)))
        self._h = h
        self.desc = krr_scene_desc()
        _check(self._lib.krr_scene_get_desc(self._h, C.byref(self.desc)))

    @property
    def n_triangles(self):
        return int(self.desc.n_triangles)

    @property
    def n_views(self):
        return int(self._lib.krr_scene_view_count(self._h))

    def camera(self, view=0):
        cam = krr_camera()
        _check(self._lib.krr_scene_get_camera(self._h, view, C.byref(cam)))
        return cam

    def view_size(self, view=0):
        w, h = C.c_int32(), C.c_int32()
        _check(self._lib.krr_scene_view_size(self._h, view, C.byref(w), C.byref(h)))
        return int(w.value), int(h.value)

    def arrays(self):
        """numpy views (no copy) of positions, indices and normals."""
        import numpy as np
        d = self.desc
        pos = np.ctypeslib.as_array(d.positions, shape=(int(d.n_vertices), 3))
        idx = np.ctypeslib.as_array(d.indices, shape=(int(d.n_triangles), 3))
        nrm = np.ctypeslib.as_array(d.normals, shape=(int(d.n_vertices), 3)) if d.normals else None
        return pos, idx, nrm

    def close(self):
        if getattr(self, "_h", None):
            self._lib.krr_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Accel:
    """Acceleration structure + closest/any-hit launch (OptixBackend counterpart)."""

    def __init__(self, scene_desc, opts=None):
        self._lib = load_library()
        h = C.c_void_p()
        _check(self._lib.krr_accel_build(C.byref(scene_desc), _json_arg(opts), C.byref(h)))
        self._h = h

    def trace_closest(self, rays_ptr, n, hits_ptr, stream=0):
        _check(self._lib.krr_accel_trace_closest(self._h, C.c_void_p(rays_ptr), n, C.c_void_p(hits_ptr),
                                                 C.c_void_p(stream)))

    def trace_closest_noalloc(self, rays_ptr, n, hits_ptr, stream=0):
        """trace_closest without device allocations (fixed per-GPU scratch;
        calls that overlap on one GPU race)."""
        _check(self._lib.krr_accel_trace_closest_noalloc(self._h, C.c_void_p(rays_ptr), n, C.c_void_p(hits_ptr),
                                                         C.c_void_p(stream)))

    def trace_any(self, rays_ptr, n, visible_ptr, stream=0):
        _check(self._lib.krr_accel_trace_any(self._h, C.c_void_p(rays_ptr), n, C.c_void_p(visible_ptr),
                                             C.c_void_p(stream)))

    def info(self):
        nn, md, nt, by = C.c_int64(), C.c_int32(), C.c_int64(), C.c_int64()
        _check(self._lib.krr_accel_info(self._h, C.byref(nn), C.byref(md), C.byref(nt), C.byref(by)))
        return {"nodes": nn.value, "max_depth": md.value, "leaf_tris": nt.value, "bytes": by.value}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.krr_accel_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WavefrontPathTracer:
    """The WavefrontPathTracer render pass over the C ABI.

    Same hooks as the reference RenderPass (renderpass.h:138-200):
    set_scene, resize, begin_frame, render, end_frame.
    """

    name = "WavefrontPathTracer"

    def __init__(self, params=None):
        self._lib = load_library()
        h = C.c_void_p()
        _check(self._lib.krr_wf_create(_json_arg(params), C.byref(h)))
        self._h = h
        self.size = (0, 0)

    @classmethod
    def from_json(cls, params):
        return cls(params)

    def to_json(self):
        buf = C.create_string_buffer(4096)
        _check(self._lib.krr_wf_get_params(self._h, buf, 4096))
        return _json.loads(buf.value.decode())

    def set_scene(self, scene):
        desc = scene.desc if isinstance(scene, Scene) else scene
        _check(self._lib.krr_wf_set_scene(self._h, C.byref(desc)))

    def resize(self, width, height):
        _check(self._lib.krr_wf_resize(self._h, int(width), int(height)))
        self.size = (int(width), int(height))

    def begin_frame(self, camera, frame_index):
        _check(self._lib.krr_wf_begin_frame(self._h, C.byref(camera), int(frame_index)))

    def render(self, film_ptr, stream=0):
        _check(self._lib.krr_wf_render(self._h, C.c_void_p(film_ptr), C.c_void_p(stream)))

    def end_frame(self):
        _check(self._lib.krr_wf_end_frame(self._h))

    def set_views(self, views):
        """views: iterable of (camera, width, height, film_ptr); all are
        rendered by render_views in one persistent launch per sample."""
        views = list(views)
        arr = (krr_view * len(views))()
        for i, (cam, w, h, film) in enumerate(views):
            arr[i].camera = cam
            arr[i].width, arr[i].height = int(w), int(h)
            arr[i].film_dev = film
        _check(self._lib.krr_wf_set_views(self._h, arr, len(views)))

    def render_views(self, frame_index, stream=0):
        _check(self._lib.krr_wf_render_views(self._h, int(frame_index), C.c_void_p(stream)))

    def autotune(self, trial_paths=0, stream=0):
        """Time persistent-launch shapes and keep the fastest; returns the report."""
        buf = C.create_string_buffer(1 << 14)
        _check(self._lib.krr_wf_autotune(self._h, int(trial_paths), buf, 1 << 14, C.c_void_p(stream)))
        return _json.loads(buf.value.decode())

    def stats(self):
        s = krr_stats()
        _check(self._lib.krr_wf_stats(self._h, C.byref(s)))
        return {f: (float(getattr(s, f)) if t is C.c_double else int(getattr(s, f))) for f, t in krr_stats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.krr_wf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MegakernelPathTracer:
    """The MegakernelPathTracer render pass over the C ABI
    (src/render/megakernel/pathtracer.h:14-50): one lane per pixel runs each
    sample's whole path, power-heuristic MIS, no media.  Same hooks as
    WavefrontPathTracer; end_frame does nothing."""

    name = "MegakernelPathTracer"

    def __init__(self, params=None):
        self._lib = load_library()
        h = C.c_void_p()
        _check(self._lib.krr_mk_create(_json_arg(params), C.byref(h)))
        self._h = h
        self.size = (0, 0)

    @classmethod
    def from_json(cls, params):
        return cls(params)

    def to_json(self):
        buf = C.create_string_buffer(4096)
        _check(self._lib.krr_mk_get_params(self._h, buf, 4096))
        return _json.loads(buf.value.decode())

    def set_scene(self, scene):
        desc = scene.desc if isinstance(scene, Scene) else scene
        _check(self._lib.krr_mk_set_scene(self._h, C.byref(desc)))

    def resize(self, width, height):
        _check(self._lib.krr_mk_resize(self._h, int(width), int(height)))
        self.size = (int(width), int(height))

    def begin_frame(self, camera, frame_index):
        _check(self._lib.krr_mk_begin_frame(self._h, C.byref(camera), int(frame_index)))

    def render(self, film_ptr, stream=0):
        _check(self._lib.krr_mk_render(self._h, C.c_void_p(film_ptr), C.c_void_p(stream)))

    def end_frame(self):
        pass

    def stats(self):
        s = krr_stats()
        _check(self._lib.krr_mk_stats(self._h, C.byref(s)))
        return {f: (float(getattr(s, f)) if t is C.c_double else int(getattr(s, f))) for f, t in krr_stats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.krr_mk_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RenderApp:
    """Headless counterpart of the reference's frame loop
    (src/core/window.cpp:412-434,457, src/main/renderer.cpp:84-97): the frame
    index starts at 1 and each frame calls beginFrame, render and endFrame of
    every pass.  The film is a device float4 buffer owned by the app.  As
    RenderApp::backBufferResized (renderer.cpp:25-31), the camera's aspect
    ratio is width / height."""

    def __init__(self, scene, passes, width, height, stream=0):
        from . import hip
        self.scene = scene
        self.passes = passes
        self.frame_index = 0
        self.width, self.height = int(width), int(height)
        self.stream = stream
        self.film = hip.DeviceBuffer(self.width * self.height * 16)
        for p in passes:
            p.set_scene(scene)
            p.resize(width, height)

    def frame(self, view=0):
        self.frame_index += 1
        cam = self.scene.camera(view)
        cam.aspect_ratio = self.width / self.height
        for p in self.passes:
            p.begin_frame(cam, self.frame_index)
        for p in self.passes:
            p.render(self.film.ptr, self.stream)
        for p in self.passes:
            p.end_frame()
        return self.film

    def film_numpy(self):
        import numpy as np
        from . import hip
        hip.synchronize()
        return self.film.to_numpy(np.float32, (self.height, self.width, 4))


def accumulate(accum_ptr, frame_ptr, n_pixels, accum_count, max_count=0, mode=0, precision=32, stream=0):
    """AccumulatePass: fold a float4 frame into the accumulation buffer and
    rewrite the frame with the mean (mode 0 sum, 1 moving average)."""
    _check(load_library().krr_accumulate(C.c_void_p(accum_ptr), C.c_void_p(frame_ptr), int(n_pixels), int(accum_count),
                                         int(max_count), int(mode), int(precision), C.c_void_p(stream)))


def error_metric(metric, img_ptr, ref_ptr, n_pixels, stream=0):
    """relMSE / MSE / MAPE / SMAPE of two device float4 images."""
    lib = load_library()
    out = C.c_double()
    _check(lib.krr_error_metric(int(metric), C.c_void_p(img_ptr), C.c_void_p(ref_ptr), int(n_pixels), C.byref(out),
                                C.c_void_p(stream)))
    return out.value
