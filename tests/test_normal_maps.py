"""Per-vertex tangents and normal maps (prepareSurfaceInteraction,
src/render/shading.h:140-145,187-203): the shading tangent is the mesh's
interpolated tangent made orthogonal to n, and a normal map tilts the frame
to n = normalize(T x + B y + N z), (x, y, z) = 2 rgb(uv) - 1, B = n x T.

Closed form that pins the checker: a Lambertian floor (y = 0, n = +y) seen
through a narrow camera and lit by a small light in direction
l = (sin b, cos b, 0) is, with direct light only, proportional to the
clamped cosine between its shading normal and l (emission, solid angle and
f = albedo / pi are the same with and without the map).  A constant map
tilting the normal by a about T gives n = T sin a + N cos a, so the image
ratio map / no map is cos(a - b) / cos b for T = +x and cos a for T = +z; a
tangent given off the plane is made orthogonal to n first.  A swapped T and
B (ratio cos a for T = +x), a sign slip in 2 rgb - 1 or a tangent left
unorthogonalised would each give another ratio.  The flat floor itself
matches the small-light closed form albedo / pi * Le * area * cos b / d^2.
"""
import math

import numpy as np
import pytest

ALPHA, BETA = math.radians(40.0), math.radians(30.0)


def _normal_texture(tmp_path, alpha=ALPHA):
    rgb = np.array([(math.sin(alpha) + 1) / 2, 0.5, (math.cos(alpha) + 1) / 2], np.float32)
    np.save(tmp_path / "nm.npy", np.broadcast_to(rgb, (2, 2, 3)).copy())
    return str(tmp_path / "nm.npy")


def _light(d=1.0, half=0.01, le=2000.0):
    lx, ly = math.sin(BETA), math.cos(BETA)
    c = np.array([d * lx, d * ly, 0.0])
    a = np.array([0.0, 0.0, 1.0])
    b = np.array([ly, -lx, 0.0])  # in the plane of x and y, perpendicular to l
    pts = [c - half * a - half * b, c + half * a - half * b, c + half * a + half * b, c - half * a + half * b]
    n = [-lx, -ly, 0.0]
    return {"type": "mesh", "material": 1, "indices": [0, 1, 2, 0, 2, 3], "emission": [le, le, le],
            "positions": [float(x) for p in pts for x in p], "normals": n * 4}


def _scene(krr, tmp_path, normal_map, tangent=None):
    floor = {"type": "mesh", "material": 0, "indices": [0, 2, 1, 0, 3, 2],
             "positions": [-1, 0, -1, 1, 0, -1, 1, 0, 1, -1, 0, 1], "normals": [0, 1, 0] * 4,
             "texcoords": [0, 0, 1, 0, 1, 1, 0, 1]}
    if tangent is not None:
        floor["tangents"] = list(tangent) * 4
    params = {"diffuse": [0.8, 0.8, 0.8]}
    if normal_map:
        params["normal_texture"] = _normal_texture(tmp_path)
    cfg = {"materials": [{"bsdf": "diffuse", "params": params}, {"bsdf": "diffuse", "params": {"diffuse": [0, 0, 0]}}],
           "model": [floor, _light()], "resolution": [16, 16],
           "camera": {"eye": [0, 1.2, 0.9], "target": [0, 0, 0], "focalLength": 1500}}
    return krr.Scene(cfg)


def _mean(oracle, s, integrator="wavefront"):
    render = oracle.render if integrator == "wavefront" else oracle.render_megakernel
    img, _ = render(s.desc, s.camera(0), 16, 16, 1, spp=512, max_depth=1, nee=1, rr=1.0, precision=64)
    return float(img[..., :3].astype(np.float64).mean())


@pytest.mark.parametrize("integrator", ["wavefront", "megakernel"])
@pytest.mark.parametrize("tangent,expected", [((1, 0, 0), math.cos(ALPHA - BETA) / math.cos(BETA)),
                                              ((0, 0, 1), math.cos(ALPHA)),
                                              ((1, 0.5, 0), math.cos(ALPHA - BETA) / math.cos(BETA))])
def test_normal_map_cosine_ratio(oracle, krr, tmp_path, integrator, tangent, expected):
    flat = _mean(oracle, _scene(krr, tmp_path, False, tangent), integrator)
    tilted = _mean(oracle, _scene(krr, tmp_path, True, tangent), integrator)
    print(f"tangent {tangent}: flat {flat:.5f} tilted {tilted:.5f} ratio {tilted / flat:.4f} expected {expected:.4f}")
    assert flat > 1e-3
    assert abs(tilted / flat - expected) < 0.01 * expected


def test_tangent_frame_does_not_change_a_lambertian_floor(oracle, krr, tmp_path):
    """Without a normal map the tangent only turns the sampling frame about
    n: the Lambertian floor's expected radiance is the same for tangents
    +x and +z and without tangents (getPerpendicular(n))."""
    vals = [_mean(oracle, _scene(krr, tmp_path, False, t)) for t in ((1, 0, 0), (0, 0, 1), None)]
    print(vals)
    assert max(vals) - min(vals) < 0.01 * vals[0]
    # and the flat floor is the small-light closed form albedo / pi * Le * area * cos b / d^2
    expected = 0.8 / math.pi * 2000.0 * 0.02 ** 2 * math.cos(BETA)
    assert abs(vals[0] - expected) < 0.005 * expected


def test_normal_map_needs_texture_coordinates(oracle, krr, tmp_path):
    """shading.h:197: the map applies only on meshes with texture
    coordinates; without them the floor renders as without the map."""
    a = _scene(krr, tmp_path, True, (1, 0, 0))
    cfg_b = _scene_cfg_without_uv(tmp_path)
    b = krr.Scene(cfg_b)
    c = _scene(krr, tmp_path, False, (1, 0, 0))
    ib, _ = oracle.render(b.desc, b.camera(0), 16, 16, 1, spp=8, max_depth=1, precision=64)
    ic, _ = oracle.render(c.desc, c.camera(0), 16, 16, 1, spp=8, max_depth=1, precision=64)
    ia, _ = oracle.render(a.desc, a.camera(0), 16, 16, 1, spp=8, max_depth=1, precision=64)
    assert np.array_equal(ib, ic)
    assert not np.array_equal(ia, ic)


def _scene_cfg_without_uv(tmp_path):
    floor = {"type": "mesh", "material": 0, "indices": [0, 2, 1, 0, 3, 2],
             "positions": [-1, 0, -1, 1, 0, -1, 1, 0, 1, -1, 0, 1], "normals": [0, 1, 0] * 4, "tangents": [1, 0, 0] * 4}
    return {"materials": [{"bsdf": "diffuse", "params": {"diffuse": [0.8, 0.8, 0.8], "normal_texture": _normal_texture(tmp_path)}},
                          {"bsdf": "diffuse", "params": {"diffuse": [0, 0, 0]}}],
            "model": [floor, _light()], "resolution": [16, 16],
            "camera": {"eye": [0, 1.2, 0.9], "target": [0, 0, 0], "focalLength": 1500}}


def test_obj_normal_map_and_tangents(krr, tmp_path):
    """OBJ import: "norm" / "map_Bump" / "disp" in the MTL become the normal
    texture (norm first, as the reference's texture slot order has it), and
    meshes with vt get per-vertex tangents: unit length, orthogonal to the
    vertex normal, pointing along increasing u."""
    np.save(tmp_path / "n.npy", np.full((2, 2, 3), 0.5, np.float32))
    np.save(tmp_path / "b.npy", np.full((2, 2, 3), 0.25, np.float32))
    (tmp_path / "m.mtl").write_text("newmtl a\nKd 0.5 0.5 0.5\nmap_Bump b.npy\nnorm n.npy\n")
    # a quad in the plane z = 0 with u along -y and v along +x (vt v is flipped on import)
    (tmp_path / "q.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 0 -1 0\nv 1 -1 0\nv 1 0 0\n"
                                    "vt 0 1\nvt 1 1\nvt 1 0\nvt 0 0\nvn 0 0 1\nusemtl a\nf 1/1/1 2/2/1 3/3/1 4/4/1\n")
    s = krr.Scene({"model": [{"model": str(tmp_path / "q.obj")}], "resolution": [8, 8],
                   "camera": {"eye": [0.5, -0.5, 2], "target": [0.5, -0.5, 0]}})
    d = s.desc
    assert d.materials[0].normal_texture == 1  # n.npy, not the bump map
    assert d.meshes[0].has_tangents == 1 and d.tangents
    t = np.ctypeslib.as_array(d.tangents, shape=(int(d.n_vertices), 3))
    assert np.allclose(np.linalg.norm(t, axis=1), 1.0, atol=1e-6)
    assert np.allclose(t, [0, -1, 0], atol=1e-6)
