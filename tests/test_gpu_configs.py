"""Parity at the measured workloads, not only at reduced sizes.

* Cornell box 256x256, 16 spp, depth 8, with the reference's own spp loop
  (L zeroed once per frame, integrator.cpp:205-259: legacy_spp_accumulate).
* The 1.31 M-triangle icosphere field at the full 1920x1080, 1 spp, depth 16,
  RR from depth 3, through the persistent launch (wavefront_depths 0) and
  through the per-depth queue stages (100).
* The same field inside the 256^3 value-noise medium (sigma_t max 20, albedo
  0.8, HG g 0.3) at the full 1280x720, 4 spp, depth 32.
* 256 views of 128^2..1024^2 pixels in one frame, 2 spp, depth 12: a view of
  every size equals the checker's image of that view.
* The radiance cache on the 1080p field: the frame's counters (queries,
  training vertices, Adam steps) and fewer rays than the plain tracer.

The checker renders each full frame in seconds on the host cores; the
images must agree pixel for pixel (1e-4 relative) with equal ray counts.
"""
import numpy as np
import pytest

from conftest import CORNELL, icosphere_scene
from test_gpu_render import assert_images_match, gpu_render

pytestmark = pytest.mark.gpu


def test_cornell_256_16spp_reference_spp_loop(krr, oracle, gpu):
    print("case: cornell 256x256 16spp depth 8 legacy_spp_accumulate", flush=True)
    s = krr.Scene(CORNELL)
    params = {"spp": 16, "max_depth": 8, "rr": 0.8, "legacy_spp_accumulate": True}
    g, gs = gpu_render(krr, gpu, s, 256, 256, params)
    o, os_ = oracle.render(s.desc, s.camera(0), 256, 256, 1, spp=16, max_depth=8, rr=0.8, legacy_spp_accumulate=1)
    print("gpu", gs["closest_rays"], gs["shadow_rays"], "cpu", os_["closest_rays"], os_["shadow_rays"], flush=True)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)
    # later samples re-add earlier samples' radiance: brighter than the mean
    m, _ = oracle.render(s.desc, s.camera(0), 256, 256, 1, spp=16, max_depth=8, rr=0.8)
    assert o[..., :3].mean() > 2.0 * m[..., :3].mean()


@pytest.fixture(scope="module")
def field_1080p(krr, oracle):
    s = krr.Scene(icosphere_scene(res=(1920, 1080)))
    o, os_ = oracle.render(s.desc, s.camera(0), 1920, 1080, 1, spp=1, max_depth=16, rr=0.8, rr_min_depth=3)
    print(f"checker 1080p frame: {os_['seconds']:.2f}s on {os_['threads']} threads", flush=True)
    return s, o, os_


@pytest.mark.parametrize("wd", [0, 100])
def test_icosphere_field_1080p_full_frame(krr, gpu, field_1080p, wd):
    print(f"case: icosphere field 1920x1080 1spp depth 16 rr3 wavefront_depths={wd}", flush=True)
    s, o, os_ = field_1080p
    params = {"spp": 1, "max_depth": 16, "rr": 0.8, "rr_min_depth": 3, "wavefront_depths": wd}
    g, gs = gpu_render(krr, gpu, s, 1920, 1080, params)
    print("gpu", gs["closest_rays"], gs["shadow_rays"], "cpu", os_["closest_rays"], os_["shadow_rays"], flush=True)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert gs["paths"] == 1920 * 1080
    assert_images_match(g, o)


def test_volume_720p_full_frame(krr, oracle, gpu):
    print("case: icosphere field + 256^3 medium 1280x720 4spp depth 32", flush=True)
    s = krr.Scene(icosphere_scene(res=(1280, 720), media=True, grid_res=256))
    params = {"spp": 4, "max_depth": 32, "rr": 0.8, "rr_min_depth": 3}
    g, gs = gpu_render(krr, gpu, s, 1280, 720, params)
    o, os_ = oracle.render(s.desc, s.camera(0), 1280, 720, 1, spp=4, max_depth=32, rr=0.8, rr_min_depth=3)
    print(f"checker: {os_['seconds']:.2f}s on {os_['threads']} threads", flush=True)
    print("gpu", gs["closest_rays"], gs["shadow_rays"], "cpu", os_["closest_rays"], os_["shadow_rays"], flush=True)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert gs["paths"] == 1280 * 720 * 4
    assert_images_match(g, o)
    # the medium scatters: more closest rays than camera paths
    assert gs["closest_rays"] > 1.2 * gs["paths"]


MULTIVIEW_256 = {"model": [{"type": "icosphere_field", "count": 1024, "level": 3, "seed": 7}],
                 "views": {"count": 256, "seed": 5, "radius": 2.0, "center": [0.5, 0.5, 0.5],
                           "sizes": [128, 256, 512, 1024]}}


def test_256_views_match_checker(krr, oracle, gpu):
    print("case: 256 views 128..1024 2spp depth 12", flush=True)
    s = krr.Scene(MULTIVIEW_256)
    n = s.n_views
    sizes = [s.view_size(v) for v in range(n)]
    assert n == 256
    params = {"spp": 2, "max_depth": 12, "rr": 0.8, "rr_min_depth": 3}
    wf = krr.WavefrontPathTracer(params)
    wf.set_scene(s)
    films = [gpu.DeviceBuffer(w * h * 16) for (w, h) in sizes]
    wf.set_views([(s.camera(v), w, h, f.ptr) for v, ((w, h), f) in enumerate(zip(sizes, films))])
    wf.render_views(1, 0)
    gpu.synchronize()
    st = wf.stats()
    total = sum(w * h for w, h in sizes)
    print(f"views {n} pixels {total} rays {st['closest_rays']} + {st['shadow_rays']}", flush=True)
    assert st["paths"] == 2 * total
    # one view of every size (the first of each), and the last view
    pick = sorted({next(v for v in range(n) if sizes[v][0] == d) for d in (128, 256, 512, 1024)} | {n - 1})
    for v in pick:
        w, h = sizes[v]
        g = films[v].to_numpy(np.float32, (h, w, 4))
        o, _ = oracle.render(s.desc, s.camera(v), w, h, 1, spp=2, max_depth=12, rr=0.8, rr_min_depth=3, view_id=v)
        print(f"view {v}: {w}x{h}", flush=True)
        assert_images_match(g, o)
    wf.close()


def test_nrc_on_1080p_field(krr, gpu):
    """The radiance cache on the 1.31 M-triangle field at 1920x1080: after 8
    frames every counter moves as configured (4 Adam steps per frame, about
    65,536 training vertices, cache queries from most render paths), the
    image is finite and non-negative, and cache-terminated paths trace fewer
    rays than the plain tracer's full paths on the same frame."""
    print("case: nrc icosphere field 1920x1080, 8 frames", flush=True)
    s = krr.Scene(icosphere_scene(res=(1920, 1080)))
    base = {"spp": 1, "max_depth": 16, "rr": 0.8, "rr_min_depth": 3}
    nrc = dict(base, nrc={"train_mod": 16, "batch": 65536, "steps_per_frame": 4, "lr": 1e-3})
    wf = krr.WavefrontPathTracer(nrc)
    wf.set_scene(s)
    wf.resize(1920, 1080)
    film = gpu.DeviceBuffer(1920 * 1080 * 16)
    for f in range(8):
        wf.begin_frame(s.camera(0), 1 + f)
        wf.render(film.ptr, 0)
        wf.end_frame()
    gpu.synchronize()
    st = wf.stats()
    img = film.to_numpy(np.float32, (1080, 1920, 4))
    wf.close()
    print({k: st[k] for k in ("closest_rays", "shadow_rays", "nrc_queries", "nrc_train_vertices", "nrc_steps", "nrc_loss")},
          flush=True)
    assert st["nrc_steps"] == 8 * 4
    assert 0.5 * 65536 < st["nrc_train_vertices"] < 1.5 * 65536
    assert st["nrc_queries"] > 0.1 * 1920 * 1080
    assert np.isfinite(st["nrc_loss"]) and st["nrc_loss"] > 0
    assert np.isfinite(img).all() and (img[..., :3] >= 0).all()
    _, plain = gpu_render(krr, gpu, s, 1920, 1080, base, frame_index=8)
    print("plain rays", plain["closest_rays"], plain["shadow_rays"], flush=True)
    assert st["closest_rays"] < 0.9 * plain["closest_rays"]
