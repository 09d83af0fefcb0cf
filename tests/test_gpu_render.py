"""Image parity of the gfx950 wavefront tracer against the CPU checker.

Same scene, camera, frame index (the per-pixel PCG32 seeds) and params on
both sides.  The checker restates the reference integrator stage by stage in
IEEE binary32 with the same elementary functions, so every path takes the
same decisions: the ray counts must be equal and every pixel must agree to
1e-4 relative (the bar of the reference comparison); in practice they are
bit-identical.
"""
import numpy as np
import pytest

from conftest import CORNELL, icosphere_scene

pytestmark = pytest.mark.gpu


def gpu_render(krr, gpu, scene, w, h, params, frame_index=1, frames=1):
    wf = krr.WavefrontPathTracer(params)
    wf.set_scene(scene)
    wf.resize(w, h)
    film = gpu.DeviceBuffer(w * h * 16)
    for f in range(frames):
        wf.begin_frame(scene.camera(0), frame_index + f)
        wf.render(film.ptr, 0)
        wf.end_frame()
    st = wf.stats()
    out = film.to_numpy(np.float32, (h, w, 4))
    wf.close()
    return out, st


def assert_images_match(g, o, rtol=1e-4, atol=1e-6):
    d = np.abs(g[..., :3] - o[..., :3])
    bad = d > rtol * np.abs(o[..., :3]) + atol
    n_bad = int(bad.any(-1).sum())
    exact = int((g[..., :3] == o[..., :3]).all(-1).sum())
    print(f"pixels: {g.shape[0] * g.shape[1]} exact: {exact} outside 1e-4: {n_bad} max abs diff {d.max():.3g}", flush=True)
    assert n_bad == 0
    assert np.all(g[..., 3] == 1.0)


def test_cornell_256_16spp_parity(krr, oracle, gpu):
    """Cornell box 256x256, 16 spp, max depth 8, NEE, rr 0.8 (frame 1)."""
    print("case: cornell 256x256 16spp depth 8", flush=True)
    s = krr.Scene(CORNELL)
    params = {"spp": 16, "max_depth": 8, "nee": True, "rr": 0.8}
    g, gs = gpu_render(krr, gpu, s, 256, 256, params)
    o, os_ = oracle.render(s.desc, s.camera(0), 256, 256, 1, spp=16, max_depth=8, nee=1, rr=0.8)
    print("gpu", gs, "cpu", os_, flush=True)
    assert gs["closest_rays"] == os_["closest_rays"]
    assert gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)
    assert 0.05 < o[..., :3].mean() < 1.0


@pytest.mark.parametrize("nee,depth,clamp,legacy", [(False, 5, False, False), (True, 0, True, False), (True, 3, False, True)])
def test_cornell_param_variants(krr, oracle, gpu, nee, depth, clamp, legacy):
    print(f"case: cornell variants nee={nee} depth={depth} clamp={clamp} legacy={legacy}", flush=True)
    s = krr.Scene(CORNELL)
    params = {"spp": 2, "max_depth": depth, "nee": nee, "enable_clamp": clamp, "clamp_max": 0.5,
              "legacy_spp_accumulate": legacy}
    g, gs = gpu_render(krr, gpu, s, 96, 80, params, frame_index=7)
    o, os_ = oracle.render(s.desc, s.camera(0), 96, 80, 7, spp=2, max_depth=depth, nee=int(nee),
                           enable_clamp=int(clamp), clamp_max=0.5, legacy_spp_accumulate=int(legacy))
    assert gs["closest_rays"] == os_["closest_rays"]
    assert gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


@pytest.mark.parametrize("wd", [0, 100])
def test_icosphere_field_parity_subsample(krr, oracle, gpu, wd):
    """The 1.31 M-triangle scene (diffuse / GGX conductor / glass) at
    240x135, 1 spp, max depth 16, RR from depth 3; all depths in the
    persistent per-path launch (0) or all as wavefront stages (100)."""
    print(f"case: icosphere field 240x135 depth 16 wavefront_depths={wd}", flush=True)
    s = krr.Scene(icosphere_scene(res=(240, 135)))
    params = {"spp": 1, "max_depth": 16, "rr": 0.8, "rr_min_depth": 3, "wavefront_depths": wd}
    g, gs = gpu_render(krr, gpu, s, 240, 135, params)
    o, os_ = oracle.render(s.desc, s.camera(0), 240, 135, 1, spp=1, max_depth=16, rr=0.8, rr_min_depth=3)
    print("gpu", gs, "cpu", os_, flush=True)
    assert gs["closest_rays"] == os_["closest_rays"]
    assert gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


@pytest.mark.parametrize("wd", [0, 100])
def test_volume_parity_subsample(krr, oracle, gpu, wd):
    """Icosphere field inside a 64^3 value-noise medium (sigma_t max 20,
    albedo 0.8, HG g 0.3): delta tracking, medium NEE with ratio tracking."""
    print(f"case: volume 128x72 spp 2 depth 32 wavefront_depths={wd}", flush=True)
    s = krr.Scene(icosphere_scene(count=256, level=2, res=(128, 72), media=True, grid_res=64))
    params = {"spp": 2, "max_depth": 32, "wavefront_depths": wd}
    g, gs = gpu_render(krr, gpu, s, 128, 72, params)
    o, os_ = oracle.render(s.desc, s.camera(0), 128, 72, 1, spp=2, max_depth=32)
    print("gpu", gs, "cpu", os_, flush=True)
    assert gs["closest_rays"] == os_["closest_rays"]
    assert gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


def test_multi_frame_seeds_differ_and_match(krr, oracle, gpu):
    """Frame index feeds the seeds: frames 1..3 each match the checker."""
    print("case: frames 1..3", flush=True)
    s = krr.Scene(CORNELL)
    wf = krr.WavefrontPathTracer({"spp": 1, "max_depth": 4})
    wf.set_scene(s)
    wf.resize(64, 64)
    film = gpu.DeviceBuffer(64 * 64 * 16)
    prev = None
    for f in (1, 2, 3):
        wf.begin_frame(s.camera(0), f)
        wf.render(film.ptr, 0)
        gpu.synchronize()
        g = film.to_numpy(np.float32, (64, 64, 4))
        o, _ = oracle.render(s.desc, s.camera(0), 64, 64, f, spp=1, max_depth=4)
        assert_images_match(g, o)
        if prev is not None:
            assert not np.array_equal(prev, g)
        prev = g
    wf.close()


@pytest.mark.parametrize("wd", [0, 1, 2, 100])
def test_megakernel_switch_depth_parity(krr, oracle, gpu, wd):
    """Depths >= wavefront_depths run per path in one persistent launch; the
    per-pixel stage order is the same, so images and ray counts must not
    depend on the switch point (0 = all megakernel, 100 = all wavefront)."""
    print(f"case: wavefront_depths={wd} cornell + volume", flush=True)
    s = krr.Scene(CORNELL)
    g, gs = gpu_render(krr, gpu, s, 128, 128, {"spp": 3, "max_depth": 8, "wavefront_depths": wd})
    o, os_ = oracle.render(s.desc, s.camera(0), 128, 128, 1, spp=3, max_depth=8)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)
    v = krr.Scene(icosphere_scene(count=128, level=2, res=(64, 48), media=True, grid_res=32))
    g, gs = gpu_render(krr, gpu, v, 64, 48, {"spp": 2, "max_depth": 12, "wavefront_depths": wd, "rr_min_depth": 2})
    o, os_ = oracle.render(v.desc, v.camera(0), 64, 48, 1, spp=2, max_depth=12, rr_min_depth=2)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


@pytest.mark.parametrize("lds_entries", [0, 4])
def test_stack_overflow_render_parity(krr, oracle, gpu, lds_entries):
    """Closest, shadow and megakernel traversal with (nearly) the whole stack
    in the global overflow area: images and ray counts are unchanged."""
    print(f"case: icosphere 96x54 stack_lds_entries={lds_entries}", flush=True)
    s = krr.Scene(icosphere_scene(count=256, level=2, res=(96, 54)))
    params = {"spp": 2, "max_depth": 10, "rr_min_depth": 3, "stack_lds_entries": lds_entries, "wavefront_depths": 2}
    g, gs = gpu_render(krr, gpu, s, 96, 54, params)
    o, os_ = oracle.render(s.desc, s.camera(0), 96, 54, 1, spp=2, max_depth=10, rr_min_depth=3)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


@pytest.mark.parametrize("nee", [True, False])
def test_emissive_furnace_parity(krr, oracle, gpu, nee):
    """Closed cube of 12 emissive diffuse triangles (every wall is a light):
    light selection over many emitters, MIS at every bounce.  The checker's
    image is pinned to the closed form in test_oracle_integrator.py."""
    from test_oracle_integrator import _furnace_scene
    print(f"case: emissive furnace 48x48 nee={nee}", flush=True)
    s = _furnace_scene(krr, 0.5, 1.0)
    params = {"spp": 4, "max_depth": 6, "nee": nee, "rr_min_depth": 1}
    g, gs = gpu_render(krr, gpu, s, 48, 48, params)
    o, os_ = oracle.render(s.desc, s.camera(0), 48, 48, 1, spp=4, max_depth=6, nee=int(nee), rr_min_depth=1)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


@pytest.mark.parametrize("albedo", [0.0, 0.9])
def test_homogeneous_medium_parity(krr, oracle, gpu, albedo):
    """Camera inside a homogeneous medium (pure absorber, and a scattering one
    with HG g = 0.5) lit by a quad emitter and looking at a diffuse floor."""
    print(f"case: homogeneous medium albedo={albedo}", flush=True)
    s = krr.Scene({"materials": [{"bsdf": "diffuse", "diffuse": [0.6, 0.5, 0.4]}],
                   "media": [{"type": "homogeneous", "sigma_t": [0.7, 0.5, 0.3], "albedo": [albedo] * 3, "g": 0.5}],
                   "model": [{"type": "quad", "corners": [[-1, 1.5, -3], [1, 1.5, -3], [1, 1.5, -1], [-1, 1.5, -1]],
                              "material": 0, "emission": [4, 4, 4]},
                             {"type": "quad", "corners": [[-4, -1, 1], [4, -1, 1], [4, -1, -6], [-4, -1, -6]],
                              "material": 0}],
                   "resolution": [64, 48],
                   "camera": {"eye": [0, 0, 0], "target": [0, -0.3, -1], "medium": 0}})
    params = {"spp": 4, "max_depth": 8}
    g, gs = gpu_render(krr, gpu, s, 64, 48, params)
    o, os_ = oracle.render(s.desc, s.camera(0), 64, 48, 1, spp=4, max_depth=8)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)
    assert o[..., :3].max() > 0


def test_camera_path_order_does_not_change_images(krr, oracle, gpu):
    """From the second sample on, the persistent launch takes up the camera
    paths that hit geometry in the previous sample first and the others
    last.  A frame rendered after four others (reordered) equals the same
    frame rendered first (pixel order), and the checker's frame, with the
    same ray counts."""
    print("case: camera-path order, icosphere field 160x90 frames 1..5", flush=True)
    s = krr.Scene(icosphere_scene(count=256, level=2, res=(160, 90)))
    params = {"spp": 2, "max_depth": 10, "rr_min_depth": 3}
    fresh, fs = gpu_render(krr, gpu, s, 160, 90, params, frame_index=5)
    wf = krr.WavefrontPathTracer(params)
    wf.set_scene(s)
    wf.resize(160, 90)
    film = gpu.DeviceBuffer(160 * 90 * 16)
    for f in range(1, 6):
        wf.begin_frame(s.camera(0), f)
        wf.render(film.ptr, 0)
    st = wf.stats()
    after = film.to_numpy(np.float32, (90, 160, 4))
    wf.close()
    assert np.array_equal(fresh, after)
    assert st["closest_rays"] == fs["closest_rays"] and st["shadow_rays"] == fs["shadow_rays"]
    o, os_ = oracle.render(s.desc, s.camera(0), 160, 90, 5, spp=2, max_depth=10, rr_min_depth=3)
    assert os_["closest_rays"] == st["closest_rays"]
    assert_images_match(after, o)
