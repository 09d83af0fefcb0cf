"""Emissive textures on mesh lights (DiffuseAreaLight::L, src/core/light.h:
188-193): a light whose material has an emissive texture emits scale *
rgb(uv), scale = max(emission) (mesh.cpp:39-59); the hit's texture
coordinates for emission seen by a ray, the sampled point's (shape.h:73-78)
for light samples.

Pins:
* Closed form: camera rays on a quadrant-textured emitter facing the camera
  return scale times the quadrant value in the orientation the texture
  coordinates give (both integrators).
* Light sampling reads the texture at the sampled point: with next-event
  estimation on and off the image of a floor under a two-toned light has the
  same expectation (FP64, block means within 5 standard errors); a wrong
  texture coordinate at the sampled point biases the NEE image.
"""
import numpy as np
import pytest

from test_textures import Q, _quadrant_texture


def _emitter(krr, tex_path, scale=2.0):
    # quad z = -1 facing +z (the camera), u = (x + 2) / 4, v = (2 - y) / 4 (v = 0 at the top)
    return krr.Scene({"materials": [{"bsdf": "diffuse", "params": {"emissive_texture": str(tex_path)}}],
                      "model": [{"type": "mesh", "positions": [-2, -2, -1, 2, -2, -1, 2, 2, -1, -2, 2, -1],
                                 "normals": [0, 0, 1] * 4, "texcoords": [0, 1, 1, 1, 1, 0, 0, 0],
                                 "indices": [0, 1, 2, 0, 2, 3], "material": 0, "emission": [scale, 1.0, 0.5]}],
                      "resolution": [32, 32], "camera": {"eye": [0, 0, 0], "target": [0, 0, -1]}})


@pytest.mark.parametrize("integrator", ["wavefront", "megakernel"])
def test_emissive_texture_quadrants(oracle, krr, tmp_path, integrator):
    np.save(tmp_path / "q.npy", _quadrant_texture())
    s = _emitter(krr, tmp_path / "q.npy")
    render = oracle.render if integrator == "wavefront" else oracle.render_megakernel
    img, _ = render(s.desc, s.camera(0), 32, 32, 1, spp=2, max_depth=0)
    lo, hi = slice(0, 7), slice(25, 32)
    for (rs, cs), want in [((lo, lo), Q[0, 0]), ((lo, hi), Q[0, 1]), ((hi, lo), Q[1, 0]), ((hi, hi), Q[1, 1])]:
        assert np.allclose(img[rs, cs, :3], 2.0 * want, rtol=2e-6), want


@pytest.mark.parametrize("integrator", ["wavefront", "megakernel"])
def test_light_samples_read_the_texture(oracle, krr, tmp_path, integrator):
    tex = np.zeros((4, 8, 3), np.float32)
    tex[:, :4] = [4.0, 1.0, 0.25]  # u < 1/2 bright, the other half dark
    np.save(tmp_path / "e.npy", tex)
    light = {"type": "mesh", "positions": [-1, 1.5, -1, 1, 1.5, -1, 1, 1.5, 1, -1, 1.5, 1], "normals": [0, -1, 0] * 4,
             "texcoords": [0, 0, 1, 0, 1, 1, 0, 1], "indices": [0, 2, 1, 0, 3, 2], "material": 0, "emission": [1, 1, 1]}
    floor = {"type": "mesh", "positions": [-3, 0, -3, 3, 0, -3, 3, 0, 3, -3, 0, 3], "normals": [0, 1, 0] * 4,
             "indices": [0, 2, 1, 0, 3, 2], "material": 1}
    s = krr.Scene({"materials": [{"bsdf": "diffuse", "params": {"emissive_texture": str(tmp_path / "e.npy")}},
                                 {"bsdf": "diffuse", "params": {"diffuse": [0.7, 0.7, 0.7]}}],
                   "model": [light, floor], "resolution": [16, 16],
                   "camera": {"eye": [0, 0.6, 3.0], "target": [0, 0, 0]}})
    render = oracle.render if integrator == "wavefront" else oracle.render_megakernel
    kw = dict(spp=1024, max_depth=1, rr=1.0, precision=64)
    if integrator == "wavefront":
        kw["rr_min_depth"] = 100
    a, _ = render(s.desc, s.camera(0), 16, 16, 1, nee=1, **kw)
    b, _ = render(s.desc, s.camera(0), 16, 16, 1, nee=0, **kw)
    # floor rows only (the light itself is seen directly in the top rows, identical in both)
    ba, bb = a[8:, :, 0].reshape(4, 2, 4, 4).mean((1, 3)), b[8:, :, 0].reshape(4, 2, 4, 4).mean((1, 3))
    assert ba.min() > 0
    assert np.allclose(ba, bb, rtol=0.08), (ba, bb)
    left, right = ba[:, :2].mean(), ba[:, 2:].mean()
    assert left > 1.5 * right  # the bright half lights the -x side


def test_obj_map_ke_and_ke(krr, tmp_path):
    """MTL Ke alone emits max(Ke) * Ke; with map_Ke the mesh keeps Ke (its
    maximum is the light's scale) and the material takes the image as its
    emissive texture (assimp.cpp:48-58,168-180; mesh.cpp:39-59)."""
    from test_scene_import import _desc_arrays
    np.save(tmp_path / "e.npy", np.ones((2, 2, 3), np.float32))
    (tmp_path / "m.mtl").write_text("newmtl lamp\nKd 0 0 0\nKe 2 1 0.5\nmap_Ke e.npy\nnewmtl glow\nKd 0 0 0\nKe 2 1 0.5\n")
    (tmp_path / "m.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvn 0 0 1\n"
                                    "usemtl lamp\nf 1/1/1 2/2/1 3/3/1\nusemtl glow\nf 1/1/1 3/3/1 4/2/1\n")
    s = krr.Scene({"model": [{"model": str(tmp_path / "m.obj")}]})
    _, _, _, meshes, mats = _desc_arrays(krr, s)
    lamp, glow = meshes
    assert list(lamp.emission) == [2, 1, 0.5] and mats[lamp.material].emissive_texture > 0
    assert list(glow.emission) == [4, 2, 1] and mats[glow.material].emissive_texture == 0
