"""The GPU binned-SAH builder (kiraray_amd/csrc/bvh_build.hip) against the
host builder (bvh.cpp) and the checker's brute force.

Both builders take the same decisions in the same float arithmetic, so on
scenes without coincident centroids they build the same tree: equal node
counts, depth and leaf sizes.  Whatever the tree, the closest hit is fixed by
the hit rule (ties to the lowest primitive), so traversal of either build
must return the brute-force hit bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import CORNELL, icosphere_scene
from test_gpu_trace import _assert_same_hits, _gpu_closest, _rays_towards_triangles

pytestmark = pytest.mark.gpu


def _soup_desc(krr, pos, idx):
    """krr_scene_desc over raw arrays: one diffuse mesh (the arrays must stay alive)."""
    d = krr.krr_scene_desc()
    d.n_vertices = len(pos)
    d.positions = pos.ctypes.data_as(C.POINTER(C.c_float))
    d.n_triangles = len(idx)
    d.indices = idx.ctypes.data_as(C.POINTER(C.c_int32))
    mesh = (krr.krr_mesh * 1)()
    mesh[0].first_triangle, mesh[0].n_triangles, mesh[0].material = 0, len(idx), 0
    mesh[0].medium_inside = mesh[0].medium_outside = -1
    mat = (krr.krr_material * 1)()
    mat[0].type = 1
    mat[0].diffuse[:] = [0.5, 0.5, 0.5]
    d.n_meshes, d.meshes = 1, mesh
    d.n_materials, d.materials = 1, mat
    return d, (mesh, mat)


class _Desc:
    def __init__(self, desc):
        self.desc = desc


@pytest.mark.parametrize("case", ["cornell", "field_82k", "field_1p3m"])
def test_device_tree_equals_host_tree(krr, gpu, case):
    """Same node count, depth and leaf triangles as the host build; traversal
    of both returns identical hits for 60,000 rays."""
    print(f"case: device vs host BVH {case}", flush=True)
    if case == "cornell":
        s = krr.Scene(CORNELL)
        lo, hi = np.array([-0.9, 0.1, -0.9]), np.array([0.9, 1.9, 2.5])
    elif case == "field_82k":
        s = krr.Scene(icosphere_scene(count=256, level=2, res=(64, 64)))
        lo, hi = np.array([-0.5, 0.0, -0.5]), np.array([1.5, 1.7, 1.5])
    else:
        s = krr.Scene(icosphere_scene())
        lo, hi = np.array([-0.5, 0.0, -0.5]), np.array([1.5, 1.7, 1.5])
    pos, idx, _ = s.arrays()
    rng = np.random.default_rng(11)
    rays = _rays_towards_triangles(pos, idx, 60000, rng, lo, hi, jitter=0.02)
    rays[::3, 3] = rng.random(len(rays[::3])).astype(np.float32)
    gd, vd, infod = _gpu_closest(krr, gpu, s, rays, {"builder": "device"})
    gh, vh, infoh = _gpu_closest(krr, gpu, s, rays, {"builder": "host"})
    print("device", infod, "host", infoh, flush=True)
    assert infod["leaf_tris"] == infoh["leaf_tris"] == s.n_triangles
    assert infod["nodes"] == infoh["nodes"] and infod["max_depth"] == infoh["max_depth"]
    _assert_same_hits(gd, gh)
    assert np.array_equal(vd, vh)
    assert (gd["prim"] >= 0).mean() > 0.3


@pytest.mark.parametrize("leaf_size,node_cost", [(1, 0.25), (4, 1.0), (8, 0.1)])
def test_device_builder_options(krr, oracle, gpu, leaf_size, node_cost):
    """Leaf size and node cost reach the device builder: hits stay the brute
    force's, the tree equals the host's."""
    print(f"case: device BVH leaf {leaf_size} cost {node_cost}", flush=True)
    s = krr.Scene(icosphere_scene(count=128, level=2, res=(64, 64)))
    pos, idx, _ = s.arrays()
    rng = np.random.default_rng(5)
    rays = _rays_towards_triangles(pos, idx, 20000, rng, np.array([-0.5, 0.0, -0.5]), np.array([1.5, 1.7, 1.5]), jitter=0.02)
    opts = {"leaf_size": leaf_size, "sah_node_cost": node_cost}
    gd, _, infod = _gpu_closest(krr, gpu, s, rays, dict(opts, builder="device"))
    _, _, infoh = _gpu_closest(krr, gpu, s, rays, dict(opts, builder="host"))
    print("device", infod, "host", infoh, flush=True)
    assert infod["nodes"] == infoh["nodes"] and infod["max_depth"] == infoh["max_depth"]
    _assert_same_hits(gd, oracle.trace_closest(s.desc, rays))


def test_device_builder_coincident_centroids(krr, oracle, gpu):
    """5,000 copies of one triangle plus a jittered field: no binned split
    separates the copies, so the builder splits them at the middle and keeps
    leaves of at most 8; every ray through the copies hits the lowest index."""
    print("case: device BVH coincident centroids", flush=True)
    rng = np.random.default_rng(9)
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    copies = 5000
    pos = [np.tile(base, (copies, 1))]
    idx = [np.arange(3 * copies, dtype=np.int32).reshape(-1, 3)]
    m = 3000
    p2 = (rng.random((m, 3, 3)) * 0.05 + rng.random((m, 1, 3)) * 2 + np.array([2.0, 0, 0])).astype(np.float32)
    pos.append(p2.reshape(-1, 3))
    idx.append((3 * copies + np.arange(3 * m, dtype=np.int32)).reshape(-1, 3))
    pos = np.ascontiguousarray(np.concatenate(pos), np.float32)
    idx = np.ascontiguousarray(np.concatenate(idx), np.int32)
    desc, keep = _soup_desc(krr, pos, idx)
    scene = _Desc(desc)
    rays = _rays_towards_triangles(pos, idx, 20000, rng, np.array([-1.0, -1.0, -2.0]), np.array([4.0, 3.0, 2.0]), jitter=0.01)
    g, vis, info = _gpu_closest(krr, gpu, scene, rays, {"builder": "device"})
    print("device", info, flush=True)
    o = oracle.trace_closest(desc, rays)
    _assert_same_hits(g, o)
    assert np.array_equal(vis, oracle.trace_any(desc, rays))
    on_copies = (o["prim"] >= 0) & (o["prim"] < copies)
    assert on_copies.sum() > 1000 and np.all(o["prim"][on_copies] == 0)
    assert info["leaf_tris"] == len(idx)
    del keep


def test_render_parity_either_builder(krr, oracle, gpu):
    """The same frame with the host-built and the device-built BVH: identical
    images and ray counts (the oracle parity tests run the default, device)."""
    print("case: render host vs device BVH", flush=True)
    s = krr.Scene(icosphere_scene(count=256, level=2, res=(96, 64)))
    imgs = []
    for b in ("host", "device"):
        wf = krr.WavefrontPathTracer({"spp": 1, "max_depth": 12, "rr_min_depth": 3, "bvh_builder": b})
        wf.set_scene(s)
        wf.resize(96, 64)
        film = gpu.DeviceBuffer(96 * 64 * 16)
        wf.begin_frame(s.camera(0), 5)
        wf.render(film.ptr, 0)
        st = wf.stats()
        imgs.append((film.to_numpy(np.float32, (64, 96, 4)), st["closest_rays"], st["shadow_rays"]))
        wf.close()
    assert np.array_equal(imgs[0][0], imgs[1][0])
    assert imgs[0][1:] == imgs[1][1:]
