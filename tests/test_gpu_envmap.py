"""GPU parity with an image environment light: the wavefront pass (with and
without media) and the megakernel path tracer against the checker, on a
field of spheres under a random RGB environment image turned by a rotation."""
import numpy as np
import pytest

from conftest import icosphere_scene
from test_gpu_megakernel import mk_render
from test_gpu_render import assert_images_match, gpu_render

pytestmark = pytest.mark.gpu


def _scene(krr, tmp_path, media=False):
    rng = np.random.default_rng(5)
    np.save(tmp_path / "env.npy", (rng.random((32, 64, 3)) * 2.0).astype(np.float32))
    cfg = icosphere_scene(count=256, level=2, res=(96, 64), media=media, grid_res=32)
    cfg["environment"] = {"image": str(tmp_path / "env.npy"), "tint": [0.9, 1.0, 1.1], "scale": 0.7,
                          "rotate": [0.9238795, 0.0, 0.3826834, 0.0]}
    return krr.Scene(cfg)


@pytest.mark.parametrize("wd", [0, 100])
def test_wavefront_envmap_parity(krr, oracle, gpu, tmp_path, wd):
    print(f"case: envmap wavefront 96x64 wavefront_depths={wd}", flush=True)
    s = _scene(krr, tmp_path)
    g, gs = gpu_render(krr, gpu, s, 96, 64, {"spp": 2, "max_depth": 8, "wavefront_depths": wd}, frame_index=4)
    o, os_ = oracle.render(s.desc, s.camera(0), 96, 64, 4, spp=2, max_depth=8)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)
    assert o[..., :3].mean() > 0.05


def test_wavefront_envmap_media_parity(krr, oracle, gpu, tmp_path):
    print("case: envmap + value-noise medium 96x64", flush=True)
    s = _scene(krr, tmp_path, media=True)
    g, gs = gpu_render(krr, gpu, s, 96, 64, {"spp": 2, "max_depth": 16}, frame_index=2)
    o, os_ = oracle.render(s.desc, s.camera(0), 96, 64, 2, spp=2, max_depth=16)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


def test_megakernel_envmap_parity(krr, oracle, gpu, tmp_path):
    print("case: envmap megakernel 96x64", flush=True)
    s = _scene(krr, tmp_path)
    g, gs = mk_render(krr, gpu, s, 96, 64, {"spp": 2, "max_depth": 8}, frame_index=4)
    o, os_ = oracle.render_megakernel(s.desc, s.camera(0), 96, 64, 4, spp=2, max_depth=8)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)
