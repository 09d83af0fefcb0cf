"""Material textures: texture coordinates from OBJ "vt" (flipped in v as
aiProcess_FlipUVs does, assimp.cpp:264) and JSON "texcoords"; map_Kd /
map_Ks (and this build's "diffuse_texture" / "specular_texture" params)
replacing the diffuse colour and the glossiness (shading.h:187-226).

Closed forms pin the checker: a plane facing the camera under a constant
sky, with BSDF sampling only, no roulette and one bounce, returns exactly
(to rounding) its texture value times the sky radiance for every path, so
the four corners of the image show the four quadrants of the texture in
the orientation the texture coordinates give them.
"""
import numpy as np
import pytest

Q = np.array([[0.2, 0.4], [0.6, 0.8]], np.float32)  # quadrant values: rows by v (top first), columns by u


def _quadrant_texture():
    t = np.repeat(np.repeat(Q, 4, 0), 4, 1)  # 8 x 8 texels
    return np.ascontiguousarray(np.repeat(t[..., None], 3, -1))


def _plane(krr, tex_path, spec_path=None, bsdf="diffuse", roughness=1.0):
    # plane z = -1 filling the view; u = (x + 2) / 4, v = (2 - y) / 4 (v = 0 at the top)
    params = {"diffuse": [1, 1, 1], "diffuse_texture": str(tex_path), "roughness": roughness}
    if spec_path is not None:
        params["specular_texture"] = str(spec_path)
    return krr.Scene({"materials": [{"bsdf": bsdf, "params": params}],
                      "model": [{"type": "mesh", "positions": [-2, -2, -1, 2, -2, -1, 2, 2, -1, -2, 2, -1],
                                 "texcoords": [0, 1, 1, 1, 1, 0, 0, 0], "indices": [0, 1, 2, 0, 2, 3], "material": 0}],
                      "environment": {"tint": [1, 1, 1], "scale": 1.0},
                      "resolution": [32, 32], "camera": {"eye": [0, 0, 0], "target": [0, 0, -1]}})


@pytest.mark.parametrize("integrator", ["wavefront", "megakernel"])
def test_texture_quadrants_closed_form(oracle, krr, tmp_path, integrator):
    np.save(tmp_path / "q.npy", _quadrant_texture())
    s = _plane(krr, tmp_path / "q.npy")
    render = oracle.render if integrator == "wavefront" else oracle.render_megakernel
    img, _ = render(s.desc, s.camera(0), 32, 32, 1, spp=4, max_depth=1, nee=0, rr=1.0)
    c = img[..., 0]
    # film row 0 is the top of the view; columns / rows 0..6 and 25..31 lie
    # where the bilinear lookup is flat inside one quadrant
    lo, hi = slice(0, 7), slice(25, 32)
    for (rs, cs), want in [((lo, lo), Q[0, 0]), ((lo, hi), Q[0, 1]), ((hi, lo), Q[1, 0]), ((hi, hi), Q[1, 1])]:
        assert np.allclose(c[rs, cs], want, rtol=2e-6), (want, c[rs, cs].min(), c[rs, cs].max())
    assert np.array_equal(img[..., 0], img[..., 2])


def test_specular_texture_sets_roughness(oracle, krr, tmp_path):
    """A conductor whose specular texture has alpha 0.7 everywhere renders
    as the same conductor with roughness 0.3 (FP64)."""
    np.save(tmp_path / "d.npy", np.full((4, 4, 3), 0.8, np.float32))
    spec = np.zeros((4, 4, 4), np.float32)
    spec[..., 3] = 0.7
    np.save(tmp_path / "s.npy", spec)
    a = _plane(krr, tmp_path / "d.npy", tmp_path / "s.npy", bsdf="conductor")
    b = _plane(krr, tmp_path / "d.npy", None, bsdf="conductor", roughness=0.3)
    ia, _ = oracle.render(a.desc, a.camera(0), 32, 32, 1, spp=2, max_depth=2, precision=64)
    ib, _ = oracle.render(b.desc, b.camera(0), 32, 32, 1, spp=2, max_depth=2, precision=64)
    assert np.allclose(ia, ib, rtol=1e-5, atol=1e-7)
    c = _plane(krr, tmp_path / "d.npy", None, bsdf="conductor", roughness=0.6)
    ic, _ = oracle.render(c.desc, c.camera(0), 32, 32, 1, spp=2, max_depth=2, precision=64)
    assert not np.allclose(ia, ic, rtol=1e-3)


def test_obj_texcoords_and_maps(krr, tmp_path):
    """vt (v flipped), faces v/t/n and v/t, map_Kd / map_Ks with options
    before the file name; a JSON material of the same name overrides the
    parameters and keeps the file's textures."""
    np.save(tmp_path / "kd.npy", np.ones((2, 3, 3), np.float32))
    np.save(tmp_path / "ks.npy", np.ones((5, 2, 4), np.float32))
    (tmp_path / "m.mtl").write_text("newmtl tex\nKd 0.5 0.5 0.5\nmap_Kd -s 1 1 1 kd.npy\nmap_Ks ks.npy\n"
                                    "newmtl plain\nKd 0.1 0.2 0.3\n")
    (tmp_path / "m.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\n"
                                    "vt 0.25 0.125\nvt 0.75 0.5\nvt 1 1\nvn 0 0 1\n"
                                    "usemtl tex\nf 1/1/1 2/2/1 3/3/1\nusemtl plain\nf 1//1 3//1 4//1\n")
    for override in (False, True):
        cfg = {"model": [{"model": str(tmp_path / "m.obj")}]}
        if override:
            cfg["materials"] = [{"name": "tex", "bsdf": "conductor", "params": {"diffuse": [0.9, 0.9, 0.9]}}]
        s = krr.Scene(cfg)
        d = s.desc
        meshes = [d.meshes[i] for i in range(d.n_meshes)]
        mats = [d.materials[m.material] for m in meshes]
        assert meshes[0].has_texcoords == 1 and meshes[1].has_texcoords == 0
        assert d.n_textures == 2
        assert (mats[0].diffuse_texture, mats[0].specular_texture) == (1, 2)
        assert (mats[1].diffuse_texture, mats[1].specular_texture) == (0, 0)
        assert (d.textures[0].width, d.textures[0].height, d.textures[1].width, d.textures[1].height) == (3, 2, 2, 5)
        idx = np.ctypeslib.as_array(d.indices, shape=(3 * d.n_triangles,))
        uv = np.ctypeslib.as_array(d.texcoords, shape=(2 * d.n_vertices,)).reshape(-1, 2)
        assert np.allclose(uv[idx[:3]], [[0.25, 0.875], [0.75, 0.5], [1, 0]])
        if override:
            assert mats[0].type == krr._abi.KRR_MAT_CONDUCTOR and abs(mats[0].diffuse[0] - 0.9) < 1e-7
    with pytest.raises(krr.KrrError):
        krr.Scene({"model": [{"type": "mesh", "positions": [0, 0, 0, 1, 0, 0, 0, 1, 0], "indices": [0, 1, 2],
                              "texcoords": [0, 0, 1]}], "materials": [{"bsdf": "diffuse"}]})
