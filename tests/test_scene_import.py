"""KiRaRay's scene JSON in krr_scene_create (kiraray_amd/csrc/scene.cpp):
Wavefront OBJ + MTL model files with the reference importer's material
mapping (src/scene/assimp.cpp:60-65,96-224), smooth normals where the file
has none (assimp's GenSmoothNormals), model transforms and material overrides
by name (src/scene/krrscene.cpp:160-238,253-350), and the camera with its
orbit controller (src/core/camera.h:64-81, camera.cpp:48-61)."""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from conftest import CORNELL

REFERENCE_CONFIG = "/root/reference/common/configs/example_cbox.json"
MAT = [{"bsdf": "diffuse", "params": {"diffuse": [0.5, 0.5, 0.5]}}]
REFERENCE_MIST = "/root/reference/common/configs/volumetric/example_cboxmist.json"


def _desc_arrays(krr, s):
    d = s.desc
    pos, idx, nrm = s.arrays()
    meshes = [d.meshes[i] for i in range(d.n_meshes)]
    mats = [d.materials[i] for i in range(d.n_materials)]
    return pos, idx, nrm, meshes, mats


def _write(path, text):
    path.write_text(text)
    return str(path)


def test_obj_polygons_materials_transform(krr, tmp_path):
    """A quad (negative indices, v//n corners) with an emissive material and
    a triangle with a glass material (d 0), under translate + scale."""
    _write(tmp_path / "m.mtl", """
newmtl lamp
  Kd 0.5 0.5 0.5
  Ke 4 3 2
  Ns 10
newmtl glass
  Kd 1 1 1
  Ni 1.33
  d 0
""")
    obj = _write(tmp_path / "m.obj", """mtllib m.mtl
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vn 0 0 1
usemtl lamp
f -4//1 -3//1 -2//1 -1//1
v 0 0 1
v 1 0 1
v 0 1 1
usemtl glass
f 5 6 7
""")
    s = krr.Scene({"model": [{"model": obj, "translate": [1, 2, 3], "scale": [2, 2, 2]}],
                   "camera": {"eye": [0, 0, 5], "target": [0, 0, 0]}})
    pos, idx, nrm, meshes, mats = _desc_arrays(krr, s)
    assert s.n_triangles == 3 and len(meshes) == 2
    lamp, glass = meshes
    assert lamp.n_triangles == 2 and glass.n_triangles == 1
    # Ke 4 3 2: the mesh light evaluates the constant Emissive texture (Ke) times scale = max(Ke)
    # (assimp.cpp:168-180, mesh.cpp:39-59, light.h:188-193)
    assert list(lamp.emission) == [16, 12, 8] and list(glass.emission) == [0, 0, 0]
    t0 = pos[idx[0]]
    assert np.allclose(t0, [[1, 2, 3], [3, 2, 3], [3, 4, 3]])  # fan (0, 1, 2), scaled 2, translated
    m_lamp, m_glass = mats[lamp.material], mats[glass.material]
    assert m_lamp.type == krr._abi.KRR_MAT_DIFFUSE and abs(m_lamp.diffuse[0] - 0.5) < 1e-7
    assert abs(m_lamp.specular[3] - (1 - math.sqrt(2 / 12))) < 1e-6  # Ns 10 -> roughness sqrt(2/12)
    assert m_glass.type == krr._abi.KRR_MAT_DIELECTRIC and m_glass.specular_transmission == 1.0
    assert abs(m_glass.ior - 1.33) < 1e-6
    assert np.allclose(nrm[idx[0]], [0, 0, 1])


def test_obj_smooth_normals_of_a_cube(krr, tmp_path):
    """A cube without vn: every corner's normal is the normalised sum of its
    three faces' normals, (+-1, +-1, +-1) / sqrt(3)."""
    v = [(x, y, z) for x in (0, 1) for y in (0, 1) for z in (0, 1)]
    faces = [(1, 2, 4, 3), (5, 7, 8, 6), (1, 5, 6, 2), (3, 4, 8, 7), (1, 3, 7, 5), (2, 6, 8, 4)]
    text = "".join(f"v {a} {b} {c}\n" for a, b, c in v) + "".join("f " + " ".join(map(str, f)) + "\n" for f in faces)
    obj = _write(tmp_path / "cube.obj", text)
    s = krr.Scene({"model": [{"model": obj}], "camera": {"eye": [3, 3, 3], "target": [0, 0, 0]}})
    pos, idx, nrm, _, _ = _desc_arrays(krr, s)
    assert s.n_triangles == 12
    for p, n in zip(pos, nrm):
        expect = np.array([1.0 if c > 0.5 else -1.0 for c in p]) / math.sqrt(3)
        assert np.allclose(n, expect, atol=1e-6), (p, n)


def test_material_override_by_name(krr, tmp_path):
    """A "materials" entry named like a file material replaces it
    (krrscene.cpp:198-229): bsdf, roughness -> specular.a = 1 - roughness."""
    _write(tmp_path / "m.mtl", "newmtl metal\n  Kd 0.2 0.2 0.2\n")
    obj = _write(tmp_path / "m.obj", "mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl metal\nf 1 2 3\n")
    s = krr.Scene({"model": [{"model": obj}],
                   "materials": [{"name": "metal", "bsdf": "conductor", "params": {"diffuse": [0.9, 0.8, 0.7], "roughness": 0.3}}],
                   "camera": {"eye": [0, 0, 2], "target": [0, 0, 0]}})
    _, _, _, meshes, mats = _desc_arrays(krr, s)
    m = mats[meshes[0].material]
    assert m.type == krr._abi.KRR_MAT_CONDUCTOR
    assert abs(m.diffuse[1] - 0.8) < 1e-7 and abs(m.specular[3] - 0.7) < 1e-6


def test_reference_camera_and_orbit_controller(krr, tmp_path):
    """camera.mData + cameraController.mData: position = target - forward *
    radius with forward = Ry(yaw) Rx(pitch) (0, 0, -1)."""
    obj = _write(tmp_path / "t.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    yaw, pitch, radius, target = 0.4, -0.3, 3.0, [0.1, 1.0, -0.2]
    s = krr.Scene({"model": [{"model": obj}],
                   "camera": {"mData": {"aspectRatio": 1.0, "focalLength": 30.0, "focalDistance": 5.0, "lensRadius": 0.0}},
                   "cameraController": {"mData": {"pitch": pitch, "yaw": yaw, "radius": radius, "target": target}}})
    c = s.camera(0)
    M = np.array(c.transform, np.float64).reshape(3, 4)
    fwd = np.array([-math.sin(yaw) * math.cos(pitch), math.sin(pitch), -math.cos(yaw) * math.cos(pitch)])
    assert np.allclose(-M[:, 2], fwd, atol=1e-6)
    assert np.allclose(M[:, 3], np.array(target) - fwd * radius, atol=1e-5)
    assert abs(c.focal_length - 30) < 1e-6 and abs(c.aspect_ratio - 1) < 1e-7 and abs(c.focal_distance - 5) < 1e-6


def test_unsupported_entries_are_errors(krr, tmp_path):
    quad = {"type": "quad", "corners": [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], "material": 0}
    bad = [
        {"model": [{"model": str(tmp_path / "x.gltf")}]},
        {"model": [{"model": str(tmp_path / "missing.obj")}]},
        {"model": [{"model": str(tmp_path / "missing.vdb"), "params": {"g": 0.5}}]},
        {"model": [quad, {"type": "light", "params": {"type": "point", "position": [0, 1, 0]}}]},
        {"model": [quad, {"type": "light", "params": {"type": "infinite", "texture": "sky.exr"}}]},
        {"model": [quad, {"type": "medium", "params": {"type": "heterogeneous", "file": str(tmp_path / "missing.vdb")}}]},
        {"model": [quad, {"type": "medium", "params": {"type": "homogeneous", "meshes": ["nothing"]}}]},
        {"model": [dict(quad, translate=[1, 0, 0])]},
        {"model": [quad], "environment": "sky.exr"},
    ]
    krr.Scene({"model": [quad], "materials": MAT})
    for cfg in bad:
        with pytest.raises(krr.KrrError):
            krr.Scene(dict(cfg, materials=MAT))


def test_light_node_infinite(krr):
    """A light node without a type is the constant infinite light of colour
    x scale (krrscene.cpp:28-47)."""
    quad = {"type": "quad", "corners": [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], "material": 0}
    s = krr.Scene({"model": [quad, {"type": "light", "name": "Sky", "params": {"color": [1, 2, 3], "scale": 2}},
                             {"type": "light"}], "materials": MAT})
    d = s.desc
    assert d.n_infinite_lights == 2
    assert list(d.infinite_lights[0].tint) == [1, 2, 3] and d.infinite_lights[0].scale == 2
    assert list(d.infinite_lights[1].tint) == [1, 1, 1] and d.infinite_lights[1].scale == 1


def test_nested_nodes_compose_transforms(krr, tmp_path):
    """A node inside a node: the child's scale, then the parent's rotation
    (90 degrees about +y) and translation, as the scene graph composes
    them; a file path given as a bare string is a model too."""
    obj = _write(tmp_path / "t.obj", "v 1 0 0\nv 0 1 0\nv 0 0 1\nf 1 2 3\n")
    h = math.sqrt(0.5)
    s = krr.Scene({"model": [{"translate": [10, 0, 0], "rotate": [h, 0, h, 0],
                              "model": [{"scale": [2, 3, 4], "model": obj}, obj]}]})
    pos, idx, nrm, meshes, _ = _desc_arrays(krr, s)
    assert len(meshes) == 2
    # R_y(90): (x, y, z) -> (z, y, -x)
    assert np.allclose(pos[idx[0]], [[10, 0, -2], [10, 3, 0], [14, 0, 0]], atol=1e-5)
    assert np.allclose(pos[idx[1]], [[10, 0, -1], [10, 1, 0], [11, 0, 0]], atol=1e-5)
    # smooth normals follow the (non-uniformly scaled) face
    P = pos[idx[0]].astype(np.float64)
    n = np.cross(P[1] - P[0], P[2] - P[0])
    assert np.allclose(nrm[idx[0][0]], n / np.linalg.norm(n), atol=1e-5)


def test_medium_node_box_and_bindings(krr, tmp_path):
    """A homogeneous medium node with sigma_a / sigma_s and a bound becomes
    sigma_t = a + s, albedo = s / t and a box without material (a medium
    interface) with the medium inside, moved by the node's transform, its faces wound outward; a second
    medium is put inside, a third outside, meshes named in the config."""
    tri = {"type": "mesh", "name": "ball", "positions": [0, 0, 0, 1, 0, 0, 0, 1, 0], "indices": [0, 1, 2]}
    other = dict(tri, name="other")
    s = krr.Scene({"model": [tri, other,
                             {"type": "medium", "name": "Mist", "translate": [0, 5, 0],
                              "params": {"g": 0.2, "sigma_a": [0.4, 0.1, 0.3], "sigma_s": [0.4, 0.3, 0.1],
                                         "bound": [[-2, -1, -3], [2, 1, 3]]}},
                             {"type": "medium", "params": {"sigma_t": [2, 2, 2], "meshes": ["ball"]}},
                             {"type": "medium", "params": {"meshes_outside": ["other", "ball"]}}], "materials": MAT})
    pos, idx, _, meshes, mats = _desc_arrays(krr, s)
    d = s.desc
    assert d.n_media == 3
    m0 = d.media[0]
    assert np.allclose(list(m0.sigma_t), [0.8, 0.4, 0.4]) and np.allclose(list(m0.albedo), [0.5, 0.75, 0.25])
    assert abs(m0.g - 0.2) < 1e-7
    assert list(d.media[1].sigma_t) == [2, 2, 2] and list(d.media[1].albedo) == [0.5, 0.5, 0.5]
    box = meshes[2]
    assert box.n_triangles == 12 and box.material == -1
    assert box.medium_inside == 0 and box.medium_outside == -1
    P = pos[idx[box.first_triangle:box.first_triangle + 12]].astype(np.float64)
    assert np.allclose(P.min((0, 1)), [-2, 4, -3]) and np.allclose(P.max((0, 1)), [2, 6, 3])
    centre = np.array([0, 5, 0])
    n = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    assert (np.einsum("ij,ij->i", n, P.mean(1) - centre) > 0).all()
    assert abs(np.linalg.norm(n, axis=1).sum() / 2 - 2 * (4 * 2 + 2 * 6 + 4 * 6)) < 1e-3  # surface area
    assert (meshes[0].medium_inside, meshes[0].medium_outside) == (1, 2)
    assert (meshes[1].medium_inside, meshes[1].medium_outside) == (-1, 2)


def test_reference_media_entry_form(krr):
    """Top-level "media" entries in the reference's form, the type inside
    "params" (krrscene.cpp:145-157), and this build's form with the type
    beside the parameters."""
    quad = {"type": "quad", "corners": [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], "material": 0}
    s = krr.Scene({"model": [quad], "media": [{"name": "fog", "params": {"type": "homogeneous", "sigma_t": [3, 3, 3]}},
                                              {"type": "homogeneous", "albedo": [0.1, 0.2, 0.3]}], "materials": MAT})
    d = s.desc
    assert d.n_media == 2 and list(d.media[0].sigma_t) == [3, 3, 3]
    assert np.allclose(list(d.media[1].albedo), [0.1, 0.2, 0.3])


@pytest.mark.skipif(not os.path.exists(REFERENCE_MIST), reason="KiRaRay tree not present")
def test_reference_example_cboxmist_scene(krr):
    """The reference's volumetric/example_cboxmist.json scene as it is: cbox.obj
    and a mist node (sigma_a = sigma_s = 0.4, g 0.2, bound [-2, 2]^3)."""
    cfg = json.load(open(REFERENCE_MIST))["scene"]
    cfg["base_dir"] = "/root/reference"
    s = krr.Scene(cfg)
    _, _, _, meshes, mats = _desc_arrays(krr, s)
    d = s.desc
    assert s.n_triangles == 36 + 12 and d.n_media == 1
    assert np.allclose(list(d.media[0].sigma_t), [0.8] * 3) and np.allclose(list(d.media[0].albedo), [0.5] * 3)
    box = [m for m in meshes if m.medium_inside == 0]
    assert len(box) == 1 and box[0].material == -1


@pytest.mark.skipif(not os.path.exists(REFERENCE_CONFIG), reason="KiRaRay tree not present")
def test_reference_example_cbox_scene(krr):
    """The reference's example_cbox.json scene, read as it is: cbox.obj with
    its MTL (light Ke = (17, 12, 4)), the orbit camera."""
    cfg = json.load(open(REFERENCE_CONFIG))["scene"]
    cfg["base_dir"] = "/root/reference"
    s = krr.Scene(cfg)
    _, _, nrm, meshes, mats = _desc_arrays(krr, s)
    lights = [m for m in meshes if any(e > 0 for e in m.emission)]
    # Ke 17 12 4 emits max(Ke) * Ke (mesh.cpp:39-59, light.h:188-193)
    assert len(lights) == 1 and list(lights[0].emission) == [289, 204, 68]
    assert s.n_triangles == 36  # 18 quads: 5 walls, two 6-sided boxes, the light
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1, atol=1e-5)
    left = [mats[m.material] for m in meshes if abs(mats[m.material].diffuse[0] - 0.63) < 1e-6]
    assert left and abs(left[0].diffuse[1] - 0.065) < 1e-6
    c = s.camera(0)
    pos = np.array(c.transform, np.float64).reshape(3, 4)[:, 3]
    assert 2.5 < pos[2] < 3.0 and 0.9 < pos[1] < 1.2  # 3.11 in front of the box


def _flat_scene(krr):
    """The Cornell box as "mesh" models with unshared vertices and flat
    float32 normals on every mesh, so that an OBJ copy can carry exactly the
    same normals."""
    src = krr.Scene(CORNELL)
    d = src.desc
    pos, idx, _ = src.arrays()
    models, mats = [], []
    for mi in range(d.n_meshes):
        m = d.meshes[mi]
        mat = d.materials[m.material]
        mats.append({"name": f"mesh{mi}", "bsdf": "diffuse", "params": {"diffuse": [float(x) for x in mat.diffuse]}})
        tri = idx[m.first_triangle:m.first_triangle + m.n_triangles]
        P = pos[tri].reshape(-1, 3).astype(np.float32)
        n = np.cross(P[1::3] - P[0::3], P[2::3] - P[0::3])
        n = (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)
        N = np.repeat(n, 3, axis=0)
        models.append({"type": "mesh", "positions": [float(x) for x in P.ravel()], "normals": [float(x) for x in N.ravel()],
                       "indices": list(range(len(P))), "material": mi, "emission": [float(x) for x in m.emission]})
    cfg = dict(CORNELL)
    cfg["model"], cfg["materials"] = models, mats
    return cfg


def _normals_from(cfg, scene):
    """cfg's "mesh" models with the per-corner normals that scene (read back
    from the OBJ copy, which renormalises vn in float32) holds."""
    _, idx, nrm = scene.arrays()
    d = scene.desc
    out = dict(cfg)
    out["model"] = []
    for mi, m in enumerate(cfg["model"]):
        me = d.meshes[mi]
        N = nrm[idx[me.first_triangle:me.first_triangle + me.n_triangles]].reshape(-1, 3)
        out["model"].append(dict(m, normals=[float(x) for x in N.ravel()]))
    return out


def _export_obj(cfg, path):
    """"mesh" models as OBJ + MTL: one usemtl run per model, vn per corner."""
    lines, mtl, nv = [f"mtllib {path.stem}.mtl"], [], 0
    for mi, m in enumerate(cfg["model"]):
        P = np.array(m["positions"], np.float32).reshape(-1, 3)
        N = np.array(m["normals"], np.float32).reshape(-1, 3)
        for p in P:
            lines.append("v " + " ".join(repr(float(x)) for x in p))
        for n in N:
            lines.append("vn " + " ".join(repr(float(x)) for x in n))
        mat = cfg["materials"][m["material"]]["params"]["diffuse"]
        mtl.append(f"newmtl mesh{mi}\nKd {mat[0]!r} {mat[1]!r} {mat[2]!r}\nKe " + " ".join(repr(e) for e in m["emission"]) + "\n")
        # the file's light emits max(Ke) * Ke (mesh.cpp:39-59, light.h:188-193): the JSON side follows
        ke = np.array(m["emission"], np.float32)
        m["emission"] = [float(x) for x in ke * ke.max()]
        lines.append(f"usemtl mesh{mi}")
        ix = np.array(m["indices"]).reshape(-1, 3) + nv + 1
        for a, b, c in ix:
            lines.append(f"f {a}//{a} {b}//{b} {c}//{c}")
        nv += len(P)
    path.write_text("\n".join(lines) + "\n")
    (path.parent / f"{path.stem}.mtl").write_text("\n".join(mtl))


@pytest.mark.gpu
def test_imported_obj_renders_like_the_same_meshes(krr, gpu, tmp_path):
    """The Cornell box given as "mesh" models and the same meshes written as
    OBJ + MTL and read back render the same image and ray counts."""
    print("case: OBJ round trip of the Cornell box", flush=True)
    cfg_a = _flat_scene(krr)
    _export_obj(cfg_a, tmp_path / "cbox.obj")
    cfg_b = dict(CORNELL)
    cfg_b["model"] = [{"model": str(tmp_path / "cbox.obj")}]
    b = krr.Scene(cfg_b)
    a = krr.Scene(_normals_from(cfg_a, b))
    assert a.n_triangles == b.n_triangles
    imgs = []
    for s in (a, b):
        wf = krr.WavefrontPathTracer({"spp": 4, "max_depth": 8})
        wf.set_scene(s)
        wf.resize(96, 96)
        film = gpu.DeviceBuffer(96 * 96 * 16)
        wf.begin_frame(a.camera(0), 2)
        wf.render(film.ptr, 0)
        st = wf.stats()
        imgs.append((film.to_numpy(np.float32, (96, 96, 4)), st["closest_rays"], st["shadow_rays"]))
        wf.close()
    assert imgs[0][1:] == imgs[1][1:]
    assert np.array_equal(imgs[0][0], imgs[1][0])
    assert imgs[0][0][..., :3].mean() > 0.05


def test_obj_copy_has_the_same_arrays(krr, tmp_path):
    """The OBJ copy of the "mesh" models reads back to the same positions,
    materials and emission bit for bit, and the same normals to float32
    rounding of their renormalisation."""
    cfg_a = _flat_scene(krr)
    _export_obj(cfg_a, tmp_path / "cbox.obj")
    cfg_b = dict(CORNELL)
    cfg_b["model"] = [{"model": str(tmp_path / "cbox.obj")}]
    a, b = krr.Scene(cfg_a), krr.Scene(cfg_b)
    pa, ia, na, ma, mta = _desc_arrays(krr, a)
    pb, ib, nb, mb, mtb = _desc_arrays(krr, b)
    assert np.array_equal(pa[ia], pb[ib])
    assert np.abs(na[ia] - nb[ib]).max() < 1e-6  # vn renormalised on import
    a2 = krr.Scene(_normals_from(cfg_a, b))
    assert np.array_equal(a2.arrays()[2][a2.arrays()[1]], nb[ib])
    assert [list(m.emission) for m in ma] == [list(m.emission) for m in mb]
    assert [list(mta[m.material].diffuse) for m in ma] == [list(mtb[m.material].diffuse) for m in mb]


@pytest.mark.gpu
def test_mist_node_render_parity(krr, oracle, gpu):
    """The Cornell box inside a mist node (sigma_a = sigma_s = 0.4, g 0.2, a
    bound box without material that the camera looks into from outside), as the
    reference's cbox mist scene sets it up: GPU and checker take the same
    paths through the box interface and the medium."""
    from test_gpu_render import assert_images_match, gpu_render
    print("case: cornell box + homogeneous mist node 96x96 spp 2", flush=True)
    cfg = dict(CORNELL)
    cfg["model"] = list(CORNELL["model"]) + [
        {"type": "medium", "name": "Mist",
         "params": {"g": 0.2, "sigma_a": [0.4, 0.4, 0.4], "sigma_s": [0.4, 0.4, 0.4], "bound": [[-2, -2, -2], [2, 2, 2]]}}]
    s = krr.Scene(cfg)
    assert s.desc.n_media == 1
    params = {"spp": 2, "max_depth": 10}
    g, gs = gpu_render(krr, gpu, s, 96, 96, params, frame_index=3)
    o, os_ = oracle.render(s.desc, s.camera(0), 96, 96, 3, spp=2, max_depth=10)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)
    clear = krr.Scene(CORNELL)  # kept alive: desc points into the scene's arrays
    plain, _ = oracle.render(clear.desc, s.camera(0), 96, 96, 3, spp=2, max_depth=10)
    assert o[..., :3].mean() < plain[..., :3].mean()  # the mist absorbs half of what it extinguishes
