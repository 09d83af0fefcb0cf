"""Image parity of the gfx950 megakernel path tracer (krr_mk_*, the
reference's MegakernelPathTracer) against the checker's restatement of it
(oracle_render_megakernel): same scene, camera, frame index and params, equal
closest and shadow ray counts and every pixel within 1e-4 relative."""
import numpy as np
import pytest

from conftest import CORNELL, icosphere_scene
from test_gpu_render import assert_images_match

pytestmark = pytest.mark.gpu


def mk_render(krr, gpu, scene, w, h, params, frame_index=1, camera=None):
    mk = krr.MegakernelPathTracer(params)
    mk.set_scene(scene)
    mk.resize(w, h)
    film = gpu.DeviceBuffer(w * h * 16)
    mk.begin_frame(camera if camera is not None else scene.camera(0), frame_index)
    mk.render(film.ptr, 0)
    st = mk.stats()
    out = film.to_numpy(np.float32, (h, w, 4))
    mk.close()
    return out, st


@pytest.mark.parametrize("nee,rr,depth,spp", [(True, 0.8, 10, 4), (False, 0.8, 8, 2), (True, 1.0, 3, 2)])
def test_cornell_parity(krr, oracle, gpu, nee, rr, depth, spp):
    print(f"case: megakernel cornell 128x96 nee={nee} rr={rr} depth={depth} spp={spp}", flush=True)
    s = krr.Scene(CORNELL)
    g, gs = mk_render(krr, gpu, s, 128, 96, {"nee": nee, "rr": rr, "max_depth": depth, "spp": spp}, frame_index=3)
    o, os_ = oracle.render_megakernel(s.desc, s.camera(0), 128, 96, 3, spp=spp, max_depth=depth, nee=int(nee), rr=rr)
    print("gpu", gs["closest_rays"], gs["shadow_rays"], "cpu", os_["closest_rays"], os_["shadow_rays"], flush=True)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert gs["paths"] == 128 * 96 * spp
    assert_images_match(g, o)
    assert o[..., :3].mean() > 0.05


def test_icosphere_field_parity(krr, oracle, gpu):
    """The 1.31 M-triangle field (diffuse, GGX conductors, glass; 16 quad
    lights) at 240 x 135, 1 spp, max depth 16."""
    print("case: megakernel icosphere field 240x135 depth 16", flush=True)
    s = krr.Scene(icosphere_scene(res=(240, 135)))
    g, gs = mk_render(krr, gpu, s, 240, 135, {"max_depth": 16, "spp": 1})
    o, os_ = oracle.render_megakernel(s.desc, s.camera(0), 240, 135, 1, spp=1, max_depth=16)
    print("gpu", gs["closest_rays"], gs["shadow_rays"], "cpu", os_["closest_rays"], os_["shadow_rays"], flush=True)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


@pytest.mark.parametrize("lds_entries", [16, 0])
def test_environment_and_stack_overflow_parity(krr, oracle, gpu, lds_entries):
    """A smaller field under a constant environment light (the miss-side MIS
    weight), with the traversal stack in LDS or wholly in the global
    overflow area."""
    print(f"case: megakernel field + environment 96x64 stack_lds_entries={lds_entries}", flush=True)
    cfg = icosphere_scene(count=256, level=2, res=(96, 64))
    cfg["environment"] = {"tint": [0.4, 0.5, 0.6], "scale": 1.0}
    s = krr.Scene(cfg)
    g, gs = mk_render(krr, gpu, s, 96, 64, {"max_depth": 8, "spp": 2, "stack_lds_entries": lds_entries}, frame_index=5)
    o, os_ = oracle.render_megakernel(s.desc, s.camera(0), 96, 64, 5, spp=2, max_depth=8)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)


def test_uneven_film_and_repeat(krr, oracle, gpu):
    """A film whose sides are not multiples of the 8 x 8 tile; the same frame
    rendered twice is bit-identical, and another frame index differs."""
    print("case: megakernel cornell 61x37 repeat", flush=True)
    s = krr.Scene(CORNELL)
    a, sa = mk_render(krr, gpu, s, 61, 37, {"spp": 2}, frame_index=7)
    b, sb = mk_render(krr, gpu, s, 61, 37, {"spp": 2}, frame_index=7)
    c, _ = mk_render(krr, gpu, s, 61, 37, {"spp": 2}, frame_index=8)
    assert np.array_equal(a, b) and sa["closest_rays"] == sb["closest_rays"]
    assert not np.array_equal(a, c)
    o, os_ = oracle.render_megakernel(s.desc, s.camera(0), 61, 37, 7, spp=2)
    assert sa["closest_rays"] == os_["closest_rays"]
    assert_images_match(a, o)
