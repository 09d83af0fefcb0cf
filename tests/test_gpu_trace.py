"""Closest/any-hit parity of the gfx950 BVH4 traversal against the CPU checker.

The hit rule (watertight test, t > 0, ties to the lowest primitive index)
makes the closest hit independent of the tree, so the GPU result must equal
the checker's brute force over all triangles bit for bit: t, primitive and
both barycentrics.
"""
import numpy as np
import pytest

from conftest import CORNELL, icosphere_scene

pytestmark = pytest.mark.gpu


def _rays_towards_triangles(pos, idx, n, rng, lo, hi, tmax=np.inf, jitter=0.2):
    tri = rng.integers(0, len(idx), n)
    b = rng.random((n, 3)).astype(np.float32)
    b /= b.sum(1, keepdims=True)
    target = (pos[idx[tri, 0]] * b[:, :1] + pos[idx[tri, 1]] * b[:, 1:2] + pos[idx[tri, 2]] * b[:, 2:3])
    org = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    d = target - org + jitter * (rng.random((n, 3)) - 0.5)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 3] = tmax
    rays[:, 4:7] = d
    return rays


def _gpu_closest(krr, gpu, scene, rays, opts=None):
    acc = krr.Accel(scene.desc, opts)
    r = gpu.DeviceBuffer.from_numpy(np.ascontiguousarray(rays, np.float32))
    hits = gpu.DeviceBuffer(len(rays) * 16)
    acc.trace_closest(r.ptr, len(rays), hits.ptr, 0)
    gpu.synchronize()
    h = hits.to_numpy(np.float32, (len(rays), 4))
    out = np.zeros(len(rays), dtype=[("t", "<f4"), ("prim", "<i4"), ("b1", "<f4"), ("b2", "<f4")])
    out["t"] = h[:, 0]
    out["prim"] = h[:, 1].view(np.int32)
    out["b1"] = h[:, 2]
    out["b2"] = h[:, 3]
    vis = gpu.DeviceBuffer(len(rays) * 4)
    acc.trace_any(r.ptr, len(rays), vis.ptr, 0)
    gpu.synchronize()
    info = acc.info()
    acc.close()
    return out, vis.to_numpy(np.int32, (len(rays),)), info


def _assert_same_hits(g, o):
    assert np.array_equal(g["prim"], o["prim"]), f"{(g['prim'] != o['prim']).sum()} primitive mismatches"
    hit = o["prim"] >= 0
    assert np.array_equal(g["t"][hit], o["t"][hit])
    assert np.array_equal(g["b1"][hit], o["b1"][hit])
    assert np.array_equal(g["b2"][hit], o["b2"][hit])


def test_cornell_closest_bit_exact(krr, oracle, gpu):
    print("case: cornell closest/any", flush=True)
    s = krr.Scene(CORNELL)
    pos, idx, _ = s.arrays()
    rng = np.random.default_rng(1)
    rays = _rays_towards_triangles(pos, idx, 20000, rng, np.array([-0.9, 0.1, -0.9]), np.array([0.9, 1.9, 2.5]))
    # a quarter with finite tmax, as shadow rays have
    rays[::4, 3] = rng.random(len(rays[::4])).astype(np.float32) * 2
    g, vis, info = _gpu_closest(krr, gpu, s, rays)
    o = oracle.trace_closest(s.desc, rays)
    _assert_same_hits(g, o)
    assert (o["prim"] >= 0).mean() > 0.5
    assert np.array_equal(vis, oracle.trace_any(s.desc, rays))
    assert info["leaf_tris"] == s.n_triangles


def test_icosphere_closest_bit_exact(krr, oracle, gpu):
    print("case: icosphere field 256 x level 2 closest/any", flush=True)
    s = krr.Scene(icosphere_scene(count=256, level=2, res=(64, 64)))
    pos, idx, _ = s.arrays()
    rng = np.random.default_rng(2)
    rays = _rays_towards_triangles(pos, idx, 30000, rng, np.array([-0.5, 0.0, -0.5]), np.array([1.5, 1.7, 1.5]), jitter=0.01)
    g, vis, info = _gpu_closest(krr, gpu, s, rays)
    o = oracle.trace_closest(s.desc, rays)
    _assert_same_hits(g, o)
    assert np.array_equal(vis, oracle.trace_any(s.desc, rays))
    assert info["max_depth"] >= 2


def test_watertight_vertices_and_edges(krr, oracle, gpu):
    """Rays aimed exactly at shared vertices and edge midpoints of closed
    icospheres from outside must never slip through."""
    print("case: watertight vertex/edge rays", flush=True)
    s = krr.Scene(icosphere_scene(count=8, level=1, res=(8, 8)))
    pos, idx, _ = s.arrays()
    sph = idx[: 8 * 80]  # the spheres come first; 80 triangles each at level 1
    targets = [pos[sph[:, 0]], 0.5 * (pos[sph[:, 0]] + pos[sph[:, 1]])]
    tg = np.concatenate(targets).astype(np.float32)
    rng = np.random.default_rng(3)
    org = np.array([0.5, 3.0, 0.5], np.float32) + rng.random((len(tg), 3)).astype(np.float32)
    rays = np.zeros((len(tg), 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 3] = np.inf
    rays[:, 4:7] = tg - org
    g, _, _ = _gpu_closest(krr, gpu, s, rays)
    o = oracle.trace_closest(s.desc, rays)
    _assert_same_hits(g, o)
    assert (o["prim"] >= 0).all(), "a ray through a shared vertex/edge missed every triangle"


def test_full_scene_bvh_matches_checker_bvh(krr, oracle, gpu):
    """1024 icospheres (1.31 M triangles): GPU hits equal the checker's own
    (independently built) BVH, itself pinned to brute force above."""
    print("case: 1.31M-triangle closest", flush=True)
    s = krr.Scene(icosphere_scene())
    assert s.n_triangles == 1024 * 1280 + 2 + 32
    pos, idx, _ = s.arrays()
    rng = np.random.default_rng(4)
    rays = _rays_towards_triangles(pos, idx, 20000, rng, np.array([-0.5, 0.05, -0.5]), np.array([1.5, 1.7, 2.5]), jitter=0.02)
    g, vis, _ = _gpu_closest(krr, gpu, s, rays)
    o = oracle.trace_closest(s.desc, rays, use_bvh=True)
    _assert_same_hits(g, o)
    assert np.array_equal(vis, oracle.trace_any(s.desc, rays, use_bvh=True))


@pytest.mark.parametrize("lds_entries", [0, 4, 96])
def test_stack_overflow_area_is_exact(krr, oracle, gpu, lds_entries):
    """The traversal stack keeps its first stack_lds_entries entries per lane
    in LDS and the rest in a global overflow area: with 0 every push goes to
    global memory, with 4 most deep pushes do, with 96 none do.  Hits must
    not change."""
    print(f"case: 1.31M-triangle closest/any, stack_lds_entries={lds_entries}", flush=True)
    s = krr.Scene(icosphere_scene())
    pos, idx, _ = s.arrays()
    rng = np.random.default_rng(5)
    rays = _rays_towards_triangles(pos, idx, 8000, rng, np.array([-0.5, 0.05, -0.5]), np.array([1.5, 1.7, 2.5]), jitter=0.05)
    g, vis, info = _gpu_closest(krr, gpu, s, rays, {"stack_lds_entries": lds_entries})
    assert 3 * info["max_depth"] + 4 > 16  # the default 16 LDS entries do not cover the whole stack
    o = oracle.trace_closest(s.desc, rays, use_bvh=True)
    _assert_same_hits(g, o)
    assert np.array_equal(vis, oracle.trace_any(s.desc, rays, use_bvh=True))


def test_repeated_calls_of_varying_size(krr, oracle, gpu):
    """The accel trace kernels are persistent (a work counter reset before
    each launch): back-to-back calls of 1, 63, 64, 65 and 5000 rays on one
    accel each trace exactly their own rays."""
    print("case: repeated accel calls n=1,63,64,65,5000", flush=True)
    s = krr.Scene(icosphere_scene(count=64, level=2, res=(16, 16)))
    pos, idx, _ = s.arrays()
    rng = np.random.default_rng(6)
    rays = _rays_towards_triangles(pos, idx, 5000, rng, np.array([-0.5, 0.05, -0.5]), np.array([1.5, 1.7, 2.5]))
    o = oracle.trace_closest(s.desc, rays)
    acc = krr.Accel(s.desc)
    r = gpu.DeviceBuffer.from_numpy(np.ascontiguousarray(rays, np.float32))
    hits = gpu.DeviceBuffer(len(rays) * 16)
    for n in (1, 63, 64, 65, 5000):
        hits.zero()
        acc.trace_closest(r.ptr, n, hits.ptr, 0)
        gpu.synchronize()
        h = hits.to_numpy(np.float32, (len(rays), 4))
        prim = h[:, 1].view(np.int32)
        assert np.array_equal(prim[:n], o["prim"][:n]), n
        assert np.array_equal(h[:n, 0][o["prim"][:n] >= 0], o["t"][:n][o["prim"][:n] >= 0]), n
        assert not h[n:].any(), f"call with n={n} wrote past its rays"
    acc.close()


@pytest.mark.parametrize("node_format", ["quantized", "float"])
def test_node_formats_on_hard_rays(krr, oracle, gpu, node_format):
    """Both node formats (128-byte float nodes, the default, and 64-byte
    quantised nodes) give the brute-force hits on rays that stress the box test:
    axis-aligned directions (zero direction components), directions with
    signed zeros and components of 1e-32 / 1e-38 (reciprocals beyond the
    slab test's 2^100 cap, or infinite), origins 10^4 away, and origins on the
    triangles' own planes (t near 0)."""
    print(f"case: hard rays, node_format={node_format}", flush=True)
    s = krr.Scene(icosphere_scene(count=64, level=2, res=(16, 16)))
    pos, idx, _ = s.arrays()
    rng = np.random.default_rng(7)
    n = 6000
    rays = np.zeros((3 * n, 8), np.float32)
    # axis-aligned: one of +-x, +-y, +-z, from points inside the scene box
    axis = rng.integers(0, 3, n)
    sign = rng.choice([-1.0, 1.0], n)
    rays[:n, 0:3] = rng.random((n, 3)) * np.array([1.0, 0.3, 1.0]) + np.array([0.0, 0.0, 0.0])
    rays[np.arange(n), 4 + axis] = sign
    # far origins aimed at triangle points
    far = _rays_towards_triangles(pos, idx, n, rng, np.array([-1e4, -1e4, -1e4]), np.array([1e4, 1e4, 1e4]), jitter=0.0)
    rays[n:2 * n] = far
    # origins on triangle points, leaving in random directions
    tri = rng.integers(0, len(idx), n)
    b = rng.random((n, 3)).astype(np.float32)
    b /= b.sum(1, keepdims=True)
    rays[2 * n:, 0:3] = pos[idx[tri, 0]] * b[:, :1] + pos[idx[tri, 1]] * b[:, 1:2] + pos[idx[tri, 2]] * b[:, 2:3]
    rays[2 * n:, 4:7] = rng.standard_normal((n, 3))
    # nearly axis-aligned: an axis direction with the other components
    # +-0.0, +-1e-32 or +-1e-38 (reciprocals above the 2^100 cap; 1e-38 is
    # subnormal and its reciprocal may overflow)
    m = n // 2
    tiny = rng.choice(np.array([0.0, -0.0, 1e-32, -1e-32, 1e-38, -1e-38], np.float32), (m, 3))
    near = np.zeros((m, 8), np.float32)
    near[:, 0:3] = rng.random((m, 3)) * np.array([1.0, 0.3, 1.0])
    near[:, 4:7] = tiny
    ax2 = rng.integers(0, 3, m)
    near[np.arange(m), 4 + ax2] = rng.choice([-1.0, 1.0], m)
    rays = np.concatenate([rays, near])
    rays[:, 3] = np.inf
    g, vis, _ = _gpu_closest(krr, gpu, s, rays, {"node_format": node_format})
    o = oracle.trace_closest(s.desc, rays)
    _assert_same_hits(g, o)
    assert np.array_equal(vis, oracle.trace_any(s.desc, rays))
    assert (o["prim"] >= 0).mean() > 0.2
