"""OpenVDB volumes in the scene importer (OpenVDBImporter::import,
src/scene/openvdb.cpp:8-67; loadOpenVDB, src/util/volume.cpp:46-78).

OpenVDB and NanoVDB are a submodule of the reference (src/ext/openvdb,
cuteday/openvdb_win) that is absent here, so the reader restates the file
format and is pinned by:

* the reference's own volume asset, common/assets/scenes/smoke/smoke.vdb
  (read where the reference tree is present): the active-voxel bounding box
  and voxel count that OpenVDB computed and stored in the grid's metadata
  (file_bbox_min/max, file_voxel_count), its transform
  (UniformScaleTranslateMap, voxel size 10/21) and the byte offsets of the
  grid, which the reader must land on exactly;
* files written by tests/vdb_writer.py from known trees (leaves, active and
  inactive tiles, float and half values, every inactive-value encoding,
  several grids per file, scale and scale-translate maps), whose dense grids
  must equal the trees' values voxel for voxel;
* the closed form of the optical depth through a constant block: voxel values
  sit at integer index points and are trilinearly interpolated with the zero
  background outside (NanoVDB SampleFromVoxels<..., 1>), and the medium's box
  is the world bounding box of index [lo, hi + 1] (NanoVDB's worldBBox), so
  an axis ray through N voxels sees (N - 1/2) voxel lengths of density.
"""
import math
import os
import struct

import numpy as np
import pytest

from vdb_writer import dense_expected, write_vdb

SMOKE = "/root/reference/common/assets/scenes/smoke/smoke.vdb"


def _scene(krr, path, params=None, node=None, extra=None):
    n = {"model": path, "params": params or {}}
    n.update(node or {})
    cfg = {"model": [n] + list(extra or []), "resolution": [8, 8], "camera": {"eye": [0, 0, 50], "target": [0, 0, 0]}}
    return krr.Scene(cfg)


def _grid(desc, i=0):
    m = desc.media[i]
    n = m.res[0] * m.res[1] * m.res[2]
    return m, np.ctypeslib.as_array(m.density, shape=(n,)).copy()


def _positions(desc):
    return np.ctypeslib.as_array(desc.positions, shape=(int(desc.n_vertices), 3)).copy()


@pytest.mark.skipif(not os.path.exists(SMOKE), reason="the reference's smoke.vdb is not on this machine")
def test_smoke_vdb_matches_its_own_metadata(krr):
    s = _scene(krr, SMOKE, {"g": 0, "sigma_a": [1, 0.5, 0.8], "sigma_s": [0.8, 0.9, 0.8]})
    d = s.desc
    m, g = _grid(d)
    # file_bbox_min (1, 2, 1), file_bbox_max (111, 223, 112), file_voxel_count 1049275
    assert list(m.res) == [111, 222, 112]
    assert int((g > 0).sum()) == 1049275
    voxel = 10.0 / 21.0
    t = np.array([-34.04761904761898, -13.095238095238095, -26.42857142857139])
    assert np.allclose(list(m.bounds_min), t + voxel * np.array([1, 2, 1]), rtol=0, atol=1e-5)
    assert np.allclose(list(m.bounds_max), t + voxel * np.array([112, 224, 113]), rtol=0, atol=1e-5)
    # sigma_t = sigma_a + sigma_s, albedo = sigma_s / sigma_t (openvdb.cpp:29-34)
    assert np.allclose(list(m.sigma_t), [1.8, 1.4, 1.6]) and np.allclose(list(m.albedo), [0.8 / 1.8, 0.9 / 1.4, 0.5])
    assert m.g == 0 and m.scale == 1
    # values are half floats; the densest voxel
    assert g.max() == np.float32(np.float16(g.max())) and 5.0 < g.max() < 6.0
    # the container: 12 triangles without material spanning the bounds, the medium inside
    assert d.n_meshes == 1 and d.n_triangles == 12
    mesh = d.meshes[0]
    assert mesh.material == -1 and mesh.medium_inside == 0 and mesh.medium_outside == -1
    p = _positions(d)
    assert np.allclose(p.min(0), list(m.bounds_min)) and np.allclose(p.max(0), list(m.bounds_max))


def _leaf(rng, frac=0.6, scale=4.0):
    vals = (np.round(rng.uniform(0, scale, 512) * 64) / 64).astype(np.float32)  # exact in half
    act = rng.uniform(0, 1, 512) < frac
    return vals, act


def _tree(rng, inactive="bg"):
    leaves = {}
    for o in [(0, 0, 0), (8, 0, 0), (-8, 16, 8), (120, -8, 0), (128, 0, -136)]:
        vals, act = _leaf(rng)
        if inactive == "bg":
            vals = np.where(act, vals, 0.0).astype(np.float32)
        elif inactive == "one":
            vals = np.where(act, vals, 0.5).astype(np.float32)
        elif inactive == "two":
            vals = np.where(act, vals, np.where(rng.uniform(0, 1, 512) < 0.5, 0.25, 1.5)).astype(np.float32)
        leaves[o] = (vals, act)
    tiles = {((16, 8, 0), 8): (2.0, True),  # an active 8^3 tile of a 16^3 node
             ((24, 8, 0), 8): (0.75, False)}  # an inactive tile holding a value
    return {"name": "density", "leaves": leaves, "tiles": tiles, "inactive": inactive}


@pytest.mark.parametrize("half", [False, True])
@pytest.mark.parametrize("compression,inactive", [(2, "bg"), (2, "one"), (2, "two"), (2, "all"), (0, "two")])
def test_written_trees_read_back_exactly(krr, tmp_path, half, compression, inactive):
    rng = np.random.default_rng(5)
    g = _tree(rng, inactive)
    g.update(half=half, voxel=(0.5, 0.25, 2.0), translate=(1.0, -2.0, 3.0))
    other = dict(_tree(np.random.default_rng(9)), name="other")
    path = write_vdb(str(tmp_path / "t.vdb"), [other, g], compression=compression)
    s = _scene(krr, path, {"sigma_t": [2, 3, 4], "albedo": [0.5, 0.25, 1], "g": 0.3, "scale": 1.5, "key_density": "density"})
    m, dense = _grid(s.desc)
    lo, hi, exp = dense_expected(g)
    assert list(m.res) == list(hi - lo + 1)
    assert np.array_equal(dense, exp)
    vox, tr = np.array([0.5, 0.25, 2.0]), np.array([1.0, -2.0, 3.0])
    assert np.allclose(list(m.bounds_min), tr + vox * lo, atol=1e-6)
    assert np.allclose(list(m.bounds_max), tr + vox * (hi + 1), atol=1e-6)
    assert list(m.sigma_t) == [2, 3, 4] and list(m.albedo) == [0.5, 0.25, 1]
    assert abs(m.g - 0.3) < 1e-7 and m.scale == 1.5
    # the first grid of the file, read by its name
    s2 = _scene(krr, path, {"key_density": "other"})
    _, d2 = _grid(s2.desc)
    assert np.array_equal(d2, dense_expected(other)[2])


@pytest.mark.parametrize("kind", ["ScaleMap", "UniformScaleMap", "UniformScaleTranslateMap"])
def test_other_maps(krr, tmp_path, kind):
    g = _tree(np.random.default_rng(1))
    g.update(map=kind, voxel=(0.5, 0.5, 0.5), translate=(3.0, 4.0, 5.0))
    path = write_vdb(str(tmp_path / "m.vdb"), [g])
    s = _scene(krr, path)  # the descriptor's arrays live as long as the scene
    m, dense = _grid(s.desc)
    lo, hi, exp = dense_expected(g)
    t = np.array([3.0, 4.0, 5.0]) if "Translate" in kind else np.zeros(3)
    assert np.allclose(list(m.bounds_min), t + 0.5 * lo) and np.array_equal(dense, exp)


def test_node_transform_places_grid_and_box(krr, tmp_path):
    g = _tree(np.random.default_rng(2))
    g.update(voxel=(0.25, 0.25, 0.25), translate=(0.0, 0.0, 0.0))
    path = write_vdb(str(tmp_path / "n.vdb"), [g])
    lo, hi, _ = dense_expected(g)
    s = _scene(krr, path, node={"translate": [10, 20, 30], "scale": [2, 3, 4]})
    m, _ = _grid(s.desc)
    sc, tr = np.array([2, 3, 4]), np.array([10, 20, 30])
    assert np.allclose(list(m.bounds_min), tr + sc * 0.25 * lo, atol=1e-5)
    assert np.allclose(list(m.bounds_max), tr + sc * 0.25 * (hi + 1), atol=1e-5)
    p = _positions(s.desc)
    assert np.allclose(p.min(0), list(m.bounds_min), atol=1e-5) and np.allclose(p.max(0), list(m.bounds_max), atol=1e-5)
    # a heterogeneous medium node naming the file is the same volume
    s2 = krr.Scene({"model": [{"type": "medium", "translate": [10, 20, 30], "scale": [2, 3, 4],
                               "params": {"type": "heterogeneous", "file": path}}],
                    "resolution": [8, 8], "camera": {"eye": [0, 0, 50], "target": [0, 0, 0]}})
    m2, _ = _grid(s2.desc)
    assert list(m2.bounds_min) == list(m.bounds_min) and list(m2.res) == list(m.res)


def _corrupt(path, offset, data):
    b = bytearray(open(path, "rb").read())
    b[offset:offset + len(data)] = data
    open(path, "wb").write(bytes(b))


@pytest.mark.parametrize("case,match", [
    ("zip", "zip or blosc"), ("blosc", "zip or blosc"), ("background", "non-zero background"),
    ("rotate", "not rotated"), ("mirror", "not rotated"), ("temperature", "temperature grid"),
    ("albedo", "albedo grid"), ("vec3", "float grids only"), ("map", "transform AffineMap"),
    ("truncated", "truncated|bad grid offsets"), ("magic", "not an OpenVDB file"), ("key", "no grid named"),
    ("instance", "instanced grid"), ("negative", "negative"), ("version", "older than 222"),
    ("far", "origin out of range")])
def test_unsupported_files_are_errors(krr, tmp_path, case, match):
    g = _tree(np.random.default_rng(3))
    grids, comp, node, params = [g], 2, {}, {}
    if case == "zip":
        comp = 3
    elif case == "blosc":
        comp = 6
    elif case == "background":
        g["background"] = 0.5
        g["inactive"] = "all"
    elif case == "rotate":
        node = {"rotate": [0.9238795, 0.0, 0.3826834, 0.0]}
    elif case == "mirror":
        node = {"scale": [-1, 1, 1]}
    elif case == "temperature":
        grids.append(dict(_tree(np.random.default_rng(4)), name="temperature"))
    elif case == "albedo":
        grids.append(dict(_tree(np.random.default_rng(4)), name="albedo"))
    elif case == "vec3":
        g["type"] = "Tree_vec3s_5_4_3"
    elif case == "map":
        g["map"] = "AffineMap"
    elif case == "key":
        params = {"key_density": "smoke"}
    elif case == "instance":
        g["instance_parent"] = "density"
    elif case == "far":  # a leaf 2^29 voxels out: origins that far would overflow the reader's ints
        g["leaves"][(1 << 29, 0, 0)] = g["leaves"][next(iter(g["leaves"]))]
    elif case == "negative":
        g["leaves"][(0, 0, 0)][0][np.nonzero(g["leaves"][(0, 0, 0)][1])[0][0]] = -1.0
    path = write_vdb(str(tmp_path / "e.vdb"), grids, compression=comp, version=221 if case == "version" else 224)
    if case == "truncated":
        b = open(path, "rb").read()
        open(path, "wb").write(b[:len(b) - 100])
    if case == "magic":
        _corrupt(path, 0, struct.pack("<q", 0x1234))
    with pytest.raises(Exception, match=match):
        _scene(krr, path, params, node=node)


def _block(n, value, voxel, origin):
    leaves = {}
    for x in range(0, n, 8):
        for y in range(0, n, 8):
            for z in range(0, n, 8):
                idx = np.arange(512)
                inside = ((x + (idx >> 6)) < n) & ((y + ((idx >> 3) & 7)) < n) & ((z + (idx & 7)) < n)
                leaves[(x, y, z)] = (np.where(inside, value, 0.0).astype(np.float32), inside)
    return {"name": "density", "leaves": leaves, "tiles": {}, "voxel": (voxel,) * 3, "translate": origin}


@pytest.mark.parametrize("n,half", [(12, False), (8, True)])
def test_constant_block_optical_depth(oracle, krr, tmp_path, n, half):
    """Absorbing block of n^3 active voxels of density rho: a ray along +x
    through the interior crosses (n - 1/2) voxel lengths of density, so the
    transmittance is exp(-sigma_t scale rho voxel (n - 1/2)) (ratio tracking
    of the CPU checker, mean of per-channel estimates, 5 standard errors).
    Voxel values at cell corners instead of index points, or the box ending
    at index hi instead of hi + 1, would give n - 1 or n lengths."""
    rho, voxel, origin, sigma, scale = 0.75, 0.1, (1.0, 2.0, 3.0), 2.0, 1.5
    g = _block(n, rho, voxel, origin)
    g["half"] = half
    path = write_vdb(str(tmp_path / "b.vdb"), [g])
    s = krr.Scene({"model": [{"model": path, "params": {"sigma_t": [sigma] * 3, "albedo": [0.5] * 3, "scale": scale}},
                             {"type": "quad", "corners": [[90, 90, 90], [91, 90, 90], [91, 91, 90], [90, 91, 90]],
                              "material": 0}],
                   "materials": [{"bsdf": "diffuse", "diffuse": [0, 0, 0]}],
                   "resolution": [8, 8], "camera": {"eye": [0, 0, 50], "target": [0, 0, 0]}})
    y, z = origin[1] + voxel * 3.3, origin[2] + voxel * (n - 2.6)
    o, dvec = (origin[0] - 0.5, y, z), (1.0, 0.0, 0.0)
    tmax = 2.0 + voxel * n
    expected = math.exp(-sigma * scale * rho * voxel * (n - 0.5))
    means, ses = [], []
    for ch in range(3):
        mu, se = oracle.transmittance(s.desc, o, dvec, tmax, -1, ch, 60000, seed=31 + ch, precision=64)
        means.append(mu)
        ses.append(se)
    mean = np.mean(means, 0)
    se = np.sqrt(np.sum(np.square(ses), 0)) / 3
    print("transmittance", mean, "expected", expected, "se", se)
    assert np.all(np.abs(mean - expected) < 5 * se + 1e-9)
