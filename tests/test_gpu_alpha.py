"""GPU parity of the any-hit alpha test (alphaKilled, src/render/shading.h:
89-113) against the checker: standalone closest / any-hit traces, the
wavefront pass (with and without a medium, whose shadow rays take the
transmittance path) and the megakernel path tracer, on scenes whose alpha
ranges over always killed, hashed and never killed."""
import numpy as np
import pytest

from conftest import CORNELL
from test_gpu_megakernel import mk_render
from test_gpu_render import assert_images_match, gpu_render
from test_gpu_trace import _assert_same_hits, _gpu_closest

pytestmark = pytest.mark.gpu


def _texture(tmp_path):
    rng = np.random.default_rng(21)
    t = (1.3 * rng.random((12, 16, 3))).astype(np.float32)  # alpha from -0.3 (killed) to 1 (kept)
    t[:, :3] = 0.0
    np.save(tmp_path / "tr.npy", t)
    return str(tmp_path / "tr.npy")


def test_trace_parity(krr, oracle, gpu, tmp_path):
    print("case: alpha-tested quads, closest and any-hit traces", flush=True)
    quad = lambda z, mat, r: {"type": "mesh", "material": mat, "indices": [0, 1, 2, 0, 2, 3],
                              "positions": [-r, -r, z, r, -r, z, r, r, z, -r, r, z],
                              "texcoords": [0, 0, 1.5, 0, 1.5, 1.25, 0, 1.25]}
    s = krr.Scene({"materials": [{"bsdf": "diffuse", "params": {"transmission_texture": _texture(tmp_path)}},
                                 {"bsdf": "diffuse", "params": {"transmission": [0.5, 0.4, 0.3]}},
                                 {"bsdf": "diffuse"}],
                   "model": [quad(-1.0, 0, 1), quad(-1.5, 1, 2), quad(-2.0, 2, 4)],
                   "resolution": [8, 8], "camera": {"eye": [0, 0, 0], "target": [0, 0, -1]}})
    rng = np.random.default_rng(5)
    n = 20000
    org = rng.uniform(-0.3, 0.3, (n, 3)).astype(np.float32)
    tgt = np.stack([rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), np.full(n, -1.0)], 1)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 3] = np.where(rng.random(n) < 0.5, np.inf, (-1.75 - org[:, 2]) / (tgt[:, 2] - org[:, 2]))
    rays[:, 4:7] = (tgt - org).astype(np.float32)
    g, gv, _ = _gpu_closest(krr, gpu, s, rays)
    o = oracle.trace_closest(s.desc, rays)
    _assert_same_hits(g, o)
    assert np.array_equal(gv, oracle.trace_any(s.desc, rays))
    prims = o["prim"]
    assert (prims < 2).any() and ((prims >= 2) & (prims < 4)).any() and (prims >= 4).any()


def _cornell(krr, tmp_path, media=False):
    quad = lambda x0, mat: {"type": "mesh", "material": mat, "indices": [0, 1, 2, 0, 2, 3],
                            "positions": [x0, 0.2, -0.3, x0 + 0.7, 0.2, -0.3, x0 + 0.7, 1.4, -0.7, x0, 1.4, -0.7],
                            "texcoords": [0, 1.5, 1.5, 1.5, 1.5, -0.25, 0, -0.25]}
    cfg = dict(CORNELL)
    cfg["materials"] = [{"bsdf": "diffuse", "params": {"diffuse": [0.8, 0.6, 0.3], "transmission_texture": _texture(tmp_path)}},
                        {"bsdf": "conductor", "params": {"diffuse": [0.9, 0.8, 0.6], "roughness": 0.3,
                                                         "transmission": [0.4, 0.5, 0.6]}}]
    cfg["model"] = list(CORNELL["model"]) + [quad(-0.85, 0), quad(0.1, 1)]
    if media:
        cfg["media"] = [{"type": "homogeneous", "sigma_t": [0.6, 0.6, 0.6], "albedo": [0.8, 0.8, 0.8]}]
        cfg["camera"] = dict(CORNELL["camera"], medium=0)
    return krr.Scene(cfg)


@pytest.mark.parametrize("media", [False, True])
def test_wavefront_alpha_parity(krr, oracle, gpu, tmp_path, media):
    print(f"case: alpha-tested quads in the Cornell box, wavefront 96x96, media={media}", flush=True)
    s = _cornell(krr, tmp_path, media)
    g, gs = gpu_render(krr, gpu, s, 96, 96, {"spp": 4, "max_depth": 8}, frame_index=3)
    o, os_ = oracle.render(s.desc, s.camera(0), 96, 96, 3, spp=4, max_depth=8)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


def test_megakernel_alpha_parity(krr, oracle, gpu, tmp_path):
    print("case: alpha-tested quads in the Cornell box, megakernel 96x96", flush=True)
    s = _cornell(krr, tmp_path)
    g, gs = mk_render(krr, gpu, s, 96, 96, {"spp": 4, "max_depth": 8}, frame_index=3)
    o, os_ = oracle.render_megakernel(s.desc, s.camera(0), 96, 96, 3, spp=4, max_depth=8)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


@pytest.mark.parametrize("integrator", ["wavefront", "megakernel"])
def test_killed_occluder_equals_plain_scene(krr, gpu, integrator):
    """The kernels with alpha tests (an always-killed quad under the light)
    render exactly what the plain kernels render without the quad."""
    print(f"case: fully killed occluder, {integrator} 64x64", flush=True)
    occ = {"type": "mesh", "material": 0, "indices": [0, 1, 2, 0, 2, 3],
           "positions": [-0.6, 1.9, -0.6, 0.6, 1.9, -0.6, 0.6, 1.9, 0.6, -0.6, 1.9, 0.6]}
    plain = krr.Scene(dict(CORNELL))
    killed = krr.Scene(dict(CORNELL, materials=[{"bsdf": "diffuse", "params": {"transmission": [2, 2, 2]}}],
                            model=list(CORNELL["model"]) + [occ]))
    render = gpu_render if integrator == "wavefront" else mk_render
    a, sa = render(krr, gpu, plain, 64, 64, {"spp": 2, "max_depth": 6}, frame_index=2)
    b, sb = render(krr, gpu, killed, 64, 64, {"spp": 2, "max_depth": 6}, frame_index=2)
    assert np.array_equal(a, b)
    assert sa["closest_rays"] == sb["closest_rays"]
