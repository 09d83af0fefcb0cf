"""The persistent launch with dynamic ray refill ("refill_lanes" > 0, the
default) and without it ("refill_lanes": 0).

A wave's traversal round ends once fewer than refill_lanes of its lanes are
still traversing; the lanes whose ray is done shade or add their light
sample and take their next ray while the others keep their traversal state.
Each path's stage order and each ray's traversal steps are unchanged, so the
images and ray counts must equal the checker's (pixels within 1e-4
relative; in practice bit-identical) for every threshold, for scenes with
and without textures and alpha tests, and must equal the launch without
refill for the radiance cache and the multi-view frame.
"""
import numpy as np
import pytest

from conftest import CORNELL, icosphere_scene
from test_gpu_alpha import _cornell
from test_gpu_render import assert_images_match, gpu_render

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("refill", [0, 1, 16, 48, 64])
def test_refill_icosphere_parity(krr, oracle, gpu, refill):
    print(f"case: refill_lanes={refill} icosphere field 240x135 depth 16", flush=True)
    s = krr.Scene(icosphere_scene(res=(240, 135)))
    params = {"spp": 1, "max_depth": 16, "rr": 0.8, "rr_min_depth": 3, "refill_lanes": refill}
    g, gs = gpu_render(krr, gpu, s, 240, 135, params)
    o, os_ = oracle.render(s.desc, s.camera(0), 240, 135, 1, spp=1, max_depth=16, rr=0.8, rr_min_depth=3)
    print("gpu", gs, "cpu", os_, flush=True)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


@pytest.mark.parametrize("refill", [16, 64])
def test_refill_cornell_spp_parity(krr, oracle, gpu, refill):
    print(f"case: refill_lanes={refill} cornell 128x128 8 spp depth 8", flush=True)
    s = krr.Scene(CORNELL)
    params = {"spp": 8, "max_depth": 8, "refill_lanes": refill}
    g, gs = gpu_render(krr, gpu, s, 128, 128, params, frame_index=5)
    o, os_ = oracle.render(s.desc, s.camera(0), 128, 128, 5, spp=8, max_depth=8)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


def test_refill_alpha_textured_parity(krr, oracle, gpu, tmp_path):
    """Textured scene with alpha-tested quads: the refill launch with any-hit
    alpha tests on closest and shadow rays."""
    print("case: refill_lanes=24 alpha-tested quads in the Cornell box 96x96", flush=True)
    s = _cornell(krr, tmp_path)
    g, gs = gpu_render(krr, gpu, s, 96, 96, {"spp": 4, "max_depth": 8, "refill_lanes": 24}, frame_index=3)
    o, os_ = oracle.render(s.desc, s.camera(0), 96, 96, 3, spp=4, max_depth=8)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


def _nrc_frames(krr, gpu, scene, refill, frames):
    params = {"spp": 1, "max_depth": 8, "refill_lanes": refill, "nrc": {"train_mod": 4, "lr": 3e-3}}
    wf = krr.WavefrontPathTracer(params)
    wf.set_scene(scene)
    wf.resize(96, 96)
    film = gpu.DeviceBuffer(96 * 96 * 16)
    stats, imgs = [], []
    for f in range(frames):
        wf.begin_frame(scene.camera(0), 1 + f)
        wf.render(film.ptr, 0)
        wf.end_frame()
        stats.append(wf.stats())
        gpu.synchronize()
        imgs.append(film.to_numpy(np.float32, (96, 96, 4)))
    wf.close()
    return imgs, stats


def test_refill_with_radiance_cache_equals_plain_launch(krr, gpu):
    """With the cache on, the first frame (queries answered by the same
    untrained network) equals the launch without refill bit for bit, and
    every frame traces the same rays and records the same queries and
    training vertices (the cache's termination rule does not read the
    network).  The trained weights are not compared: which training records
    fill which step of the batch depends on the order of the atomic pushes."""
    print("case: refill_lanes=16 vs 0 with the radiance cache, cornell 96x96, 6 frames", flush=True)
    s = krr.Scene(CORNELL)
    a, sa = _nrc_frames(krr, gpu, s, 0, 6)
    b, sb = _nrc_frames(krr, gpu, s, 16, 6)
    for x, y in zip(sa, sb):
        for k in ("closest_rays", "shadow_rays", "nrc_queries", "nrc_train_vertices", "nrc_steps"):
            assert x[k] == y[k], k
    assert sa[-1]["nrc_queries"] > 0
    assert np.array_equal(a[0], b[0])


def test_refill_views_equal_plain_launch(krr, gpu):
    print("case: refill_lanes=32 vs 0, 4 views of an icosphere field", flush=True)
    cfg = icosphere_scene(count=64, level=2, res=(64, 64))
    cfg["views"] = {"count": 4, "seed": 3, "radius": 2.0, "sizes": [24, 40, 56]}
    s = krr.Scene(cfg)
    sizes = [s.view_size(v) for v in range(s.n_views)]
    out = []
    for refill in (0, 32):
        wf = krr.WavefrontPathTracer({"spp": 2, "max_depth": 12, "rr_min_depth": 3, "refill_lanes": refill})
        wf.set_scene(s)
        films = [gpu.DeviceBuffer(w * h * 16) for (w, h) in sizes]
        wf.set_views([(s.camera(v), w, h, f.ptr) for v, ((w, h), f) in enumerate(zip(sizes, films))])
        wf.render_views(3, 0)
        st = wf.stats()
        gpu.synchronize()
        out.append(([f.to_numpy(np.float32, (h, w, 4)) for f, (w, h) in zip(films, sizes)], st))
        wf.close()
    (ia, sa), (ib, sb) = out
    assert sa["closest_rays"] == sb["closest_rays"] and sa["shadow_rays"] == sb["shadow_rays"]
    for x, y in zip(ia, ib):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("lds_nodes", [21, 341])
def test_lds_node_copy_parity(krr, oracle, gpu, lds_nodes):
    """The refill launch reading the top tree nodes from a per-block LDS copy
    ("lds_nodes") traces the same rays with the same hits."""
    print(f"case: lds_nodes={lds_nodes} icosphere field 240x135 depth 16", flush=True)
    s = krr.Scene(icosphere_scene(res=(240, 135)))
    params = {"spp": 1, "max_depth": 16, "rr": 0.8, "rr_min_depth": 3, "lds_nodes": lds_nodes}
    g, gs = gpu_render(krr, gpu, s, 240, 135, params)
    o, os_ = oracle.render(s.desc, s.camera(0), 240, 135, 1, spp=1, max_depth=16, rr=0.8, rr_min_depth=3)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


@pytest.mark.parametrize("leaf_share", [0, 30, 70])
def test_leaf_phase_parity(krr, oracle, gpu, leaf_share):
    """Iterations that take either leaves or inner nodes ("leaf_share")
    trace the same rays with the same hits."""
    print(f"case: leaf_share={leaf_share} icosphere field 240x135 depth 16", flush=True)
    s = krr.Scene(icosphere_scene(res=(240, 135)))
    params = {"spp": 1, "max_depth": 16, "rr": 0.8, "rr_min_depth": 3, "leaf_share": leaf_share}
    g, gs = gpu_render(krr, gpu, s, 240, 135, params)
    o, os_ = oracle.render(s.desc, s.camera(0), 240, 135, 1, spp=1, max_depth=16, rr=0.8, rr_min_depth=3)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


def test_refill_by_scene_size(krr, gpu):
    """refill_lanes -1 (the default): the refill launch for trees of at least
    4096 nodes, rays run to their end for smaller ones (the 36-triangle
    Cornell box); an explicit value is kept."""
    print("case: refill_lanes -1 resolved by the tree size", flush=True)
    for cfg, expected in ((CORNELL, 0), (icosphere_scene(count=64, level=2, res=(64, 64)), 8)):
        wf = krr.WavefrontPathTracer({"spp": 1})
        assert wf.to_json()["refill_lanes"] == -1
        wf.set_scene(krr.Scene(cfg))
        assert wf.to_json()["refill_lanes_in_force"] == expected
        wf.close()
    wf = krr.WavefrontPathTracer({"refill_lanes": 16})
    wf.set_scene(krr.Scene(CORNELL))
    assert wf.to_json()["refill_lanes_in_force"] == 16
    wf.close()
