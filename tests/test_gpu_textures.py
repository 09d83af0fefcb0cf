"""GPU parity with textured materials: the Cornell box with two textured
quads (a diffuse one with a random colour texture, a rough conductor whose
glossiness comes from a random specular texture), through the wavefront
pass and the megakernel path tracer, against the checker."""
import numpy as np
import pytest

from conftest import CORNELL
from test_gpu_megakernel import mk_render
from test_gpu_render import assert_images_match, gpu_render

pytestmark = pytest.mark.gpu


def _scene(krr, tmp_path):
    rng = np.random.default_rng(9)
    np.save(tmp_path / "kd.npy", rng.random((16, 24, 3), dtype=np.float32))
    spec = rng.random((8, 8, 4), dtype=np.float32)
    spec[..., 3] = 0.3 + 0.6 * spec[..., 3]
    np.save(tmp_path / "ks.npy", spec)
    quad = lambda x0, mat: {"type": "mesh", "material": mat, "indices": [0, 1, 2, 0, 2, 3],
                            "positions": [x0, 0.2, -0.5, x0 + 0.7, 0.2, -0.5, x0 + 0.7, 1.2, -0.9, x0, 1.2, -0.9],
                            "texcoords": [0, 1.5, 1.5, 1.5, 1.5, -0.25, 0, -0.25]}
    cfg = dict(CORNELL)
    cfg["materials"] = [{"bsdf": "diffuse", "params": {"diffuse_texture": str(tmp_path / "kd.npy")}},
                        {"bsdf": "conductor", "params": {"diffuse": [0.9, 0.8, 0.6], "specular_texture": str(tmp_path / "ks.npy")}}]
    cfg["model"] = list(CORNELL["model"]) + [quad(-0.85, 0), quad(0.1, 1)]
    return krr.Scene(cfg)


def test_wavefront_texture_parity(krr, oracle, gpu, tmp_path):
    print("case: textured quads in the Cornell box, wavefront 96x96", flush=True)
    s = _scene(krr, tmp_path)
    g, gs = gpu_render(krr, gpu, s, 96, 96, {"spp": 4, "max_depth": 8}, frame_index=3)
    o, os_ = oracle.render(s.desc, s.camera(0), 96, 96, 3, spp=4, max_depth=8)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


def test_megakernel_texture_parity(krr, oracle, gpu, tmp_path):
    print("case: textured quads in the Cornell box, megakernel 96x96", flush=True)
    s = _scene(krr, tmp_path)
    g, gs = mk_render(krr, gpu, s, 96, 96, {"spp": 4, "max_depth": 8}, frame_index=3)
    o, os_ = oracle.render_megakernel(s.desc, s.camera(0), 96, 96, 3, spp=4, max_depth=8)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


def _emissive_scene(krr, tmp_path):
    rng = np.random.default_rng(13)
    np.save(tmp_path / "ke.npy", rng.random((10, 14, 3), dtype=np.float32))
    lamp = {"type": "mesh", "material": 0, "indices": [0, 2, 1, 0, 3, 2], "emission": [3.0, 2.0, 1.0],
            "positions": [-0.5, 1.6, -0.6, 0.4, 1.6, -0.6, 0.4, 1.6, 0.2, -0.5, 1.6, 0.2], "normals": [0, -1, 0] * 4,
            "texcoords": [0, 0, 1.25, 0, 1.25, 1.5, 0, 1.5]}
    cfg = dict(CORNELL)
    cfg["materials"] = [{"bsdf": "diffuse", "params": {"emissive_texture": str(tmp_path / "ke.npy")}}]
    cfg["model"] = list(CORNELL["model"]) + [lamp]
    return krr.Scene(cfg)


@pytest.mark.parametrize("integrator", ["wavefront", "megakernel"])
def test_emissive_texture_parity(krr, oracle, gpu, tmp_path, integrator):
    print(f"case: emissive-textured lamp in the Cornell box, {integrator} 96x96", flush=True)
    s = _emissive_scene(krr, tmp_path)
    if integrator == "wavefront":
        g, gs = gpu_render(krr, gpu, s, 96, 96, {"spp": 4, "max_depth": 8}, frame_index=3)
        o, os_ = oracle.render(s.desc, s.camera(0), 96, 96, 3, spp=4, max_depth=8)
    else:
        g, gs = mk_render(krr, gpu, s, 96, 96, {"spp": 4, "max_depth": 8}, frame_index=3)
        o, os_ = oracle.render_megakernel(s.desc, s.camera(0), 96, 96, 3, spp=4, max_depth=8)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


def _normal_mapped_scene(krr, tmp_path):
    """The Cornell box plus a quad with per-vertex tangents, a random normal
    map (tangent-space normals tilted by up to about 25 degrees) and a
    diffuse texture."""
    rng = np.random.default_rng(17)
    nm = np.empty((12, 10, 3), np.float32)
    nm[..., :2] = 0.5 + 0.35 * (rng.random((12, 10, 2), dtype=np.float32) - 0.5) * 2
    nm[..., 2] = 0.85 + 0.15 * rng.random((12, 10), dtype=np.float32)
    np.save(tmp_path / "nm.npy", nm)
    np.save(tmp_path / "kd2.npy", rng.random((6, 6, 3), dtype=np.float32))
    quad = {"type": "mesh", "material": 2, "indices": [0, 1, 2, 0, 2, 3],
            "positions": [-0.6, 0.05, -0.3, 0.5, 0.05, -0.3, 0.5, 0.9, -0.8, -0.6, 0.9, -0.8],
            "normals": [0, 0.51, 0.86] * 4, "tangents": [1, 0.2, 0] * 2 + [0.9, 0, 0.3] * 2,
            "texcoords": [0, 0, 2, 0, 2, 1.5, 0, 1.5]}
    cfg = dict(CORNELL)
    cfg["materials"] = [{"bsdf": "diffuse", "params": {"diffuse_texture": str(tmp_path / "kd2.npy"),
                                                       "normal_texture": str(tmp_path / "nm.npy")}}]
    quad["material"] = 0
    cfg["model"] = list(CORNELL["model"]) + [quad]
    return krr.Scene(cfg)


@pytest.mark.parametrize("integrator", ["wavefront", "megakernel"])
def test_normal_map_parity(krr, oracle, gpu, tmp_path, integrator):
    print(f"case: normal-mapped quad with tangents in the Cornell box, {integrator} 96x96", flush=True)
    s = _normal_mapped_scene(krr, tmp_path)
    if integrator == "wavefront":
        g, gs = gpu_render(krr, gpu, s, 96, 96, {"spp": 4, "max_depth": 8}, frame_index=5)
        o, os_ = oracle.render(s.desc, s.camera(0), 96, 96, 5, spp=4, max_depth=8)
    else:
        g, gs = mk_render(krr, gpu, s, 96, 96, {"spp": 4, "max_depth": 8}, frame_index=5)
        o, os_ = oracle.render_megakernel(s.desc, s.camera(0), 96, 96, 5, spp=4, max_depth=8)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)
