"""The any-hit alpha test (alphaKilled, src/render/shading.h:89-113) of the
checker: a material with a transmission texture or constant T (pbrt.cpp:
145-156 stores 1 - opacity there) ignores a hit when alpha = 1 -
luminance(T) is <= 0, or is < 1 and HashFloat(origin, direction) > alpha
(hash.h:130-134, MurmurHash64A over the 24 bytes of the ray).

Pins:
* MurmurHash64A reproduces SMHasher's published verification code for it
  (0x1F0D3804, "Murmur2B"): 256 keys 0..i-1 hashed with seeds 256 - i, the
  concatenated hashes hashed with seed 0; a wrong constant, shift, block or
  tail case fails it.
* Closed forms on two parallel quads: T = 0 keeps every hit, alpha <= 0
  kills every hit (closest and shadow rays), a texture that is opaque on one
  half and clear on the other splits the hits by the texture coordinate.
* alpha = 0.2 kills 80 % of the hits (the direction of the comparison), and
  which ones follows the hash of the ray's bytes, origin first.
* An occluder whose every hit is killed changes nothing: renders with and
  without it are identical, for both integrators - closest and shadow rays
  are both tested (it hangs under the light of the Cornell box).
"""
import struct

import numpy as np
import pytest

from conftest import CORNELL

LUM = (np.float32(0.299), np.float32(0.587), np.float32(0.114))


def test_murmur64a_smhasher_verification(oracle):
    key = bytes(range(256))
    hashes = b"".join(struct.pack("<Q", oracle.murmur64a(key[:i], 256 - i)) for i in range(256))
    assert oracle.murmur64a(hashes, 0) & 0xFFFFFFFF == 0x1F0D3804


def _quads(krr, front_params):
    # front quad z = -1 (prims 0, 1) with the tested material, opaque back quad z = -2 (prims 2, 3);
    # u = (x + 1) / 2, v = (y + 1) / 2 on the front quad
    # (the back quad is 4 times larger, so every ray through the front quad meets it)
    quad = lambda z, mat, r: {"type": "mesh", "material": mat, "indices": [0, 1, 2, 0, 2, 3],
                              "positions": [-r, -r, z, r, -r, z, r, r, z, -r, r, z],
                              "texcoords": [0, 0, 1, 0, 1, 1, 0, 1]}
    return krr.Scene({"materials": [{"bsdf": "diffuse", "params": front_params}, {"bsdf": "diffuse"}],
                      "model": [quad(-1.0, 0, 1), quad(-2.0, 1, 4)],
                      "resolution": [8, 8], "camera": {"eye": [0, 0, 0], "target": [0, 0, -1]}})


def _rays(n, seed, u_range=(0.05, 0.95), tmax=np.inf):
    rng = np.random.default_rng(seed)
    u = rng.uniform(*u_range, n)
    v = rng.uniform(0.05, 0.95, n)
    org = rng.uniform(-0.2, 0.2, (n, 3)).astype(np.float32)
    org[:, 2] = rng.uniform(-0.3, 0.3, n)
    target = np.stack([2 * u - 1, 2 * v - 1, np.full(n, -1.0)], 1)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 3] = tmax
    rays[:, 4:7] = (target - org).astype(np.float32)
    return rays, u


def _alpha_of(T):
    T = np.asarray(T, np.float32)
    return np.float32(1) - (T[0] * LUM[0] + T[1] * LUM[1] + T[2] * LUM[2])


@pytest.mark.parametrize("use_bvh", [False, True])
def test_constant_transmission_closed_forms(oracle, krr, use_bvh):
    rays, _ = _rays(512, 1)
    opaque = _quads(krr, {"transmission": [0, 0, 0]})
    clear = _quads(krr, {"transmission": [2, 2, 2]})  # alpha = -1: every hit killed
    assert (oracle.trace_closest(opaque.desc, rays, use_bvh)["prim"] < 2).all()
    h = oracle.trace_closest(clear.desc, rays, use_bvh)
    assert (h["prim"] >= 2).all()
    # shadow rays that end before the back quad: blocked by the opaque front quad only
    short = rays.copy()
    short[:, 3] = (-1.5 - short[:, 2]) / short[:, 6]  # t of the plane z = -1.5, between the quads
    assert (oracle.trace_any(opaque.desc, short, use_bvh) == 0).all()
    assert (oracle.trace_any(clear.desc, short, use_bvh) == 1).all()


def test_texture_splits_hits_by_uv(oracle, krr, tmp_path):
    tex = np.zeros((2, 8, 3), np.float32)
    tex[:, 4:] = 2.0  # u > 1/2 clear; texel centres flat on [1/16, 7/16] and [9/16, 15/16]
    np.save(tmp_path / "t.npy", tex)
    s = _quads(krr, {"transmission_texture": str(tmp_path / "t.npy")})
    for lo, hi, front in [(0.1, 0.4, True), (0.6, 0.9, False)]:
        rays, _ = _rays(256, 2, (lo, hi))
        for use_bvh in (False, True):
            h = oracle.trace_closest(s.desc, rays, use_bvh)
            assert ((h["prim"] < 2) == front).all()


def test_partial_alpha_follows_the_ray_hash(oracle, krr):
    T = [0.8, 0.8, 0.8]
    alpha = _alpha_of(T)
    assert 0.19 < alpha < 0.21
    s = _quads(krr, {"transmission": T})
    rays, _ = _rays(4000, 3)
    killed = oracle.trace_closest(s.desc, rays)["prim"] >= 2
    assert abs(killed.mean() - (1 - alpha)) < 0.03
    u = np.array([oracle.murmur64a(r[0:3].tobytes() + r[4:7].tobytes()) & 0xFFFFFFFF for r in rays], np.float64)
    want = (u.astype(np.float32) * np.float32(2.0 ** -32)) > alpha
    assert np.array_equal(killed, want)


@pytest.mark.parametrize("integrator", ["wavefront", "megakernel"])
def test_fully_killed_occluder_changes_nothing(oracle, krr, integrator):
    base = dict(CORNELL)
    occ = {"type": "mesh", "material": 0, "indices": [0, 1, 2, 0, 2, 3],
           "positions": [-0.6, 1.9, -0.6, 0.6, 1.9, -0.6, 0.6, 1.9, 0.6, -0.6, 1.9, 0.6]}
    with_occ = dict(base, materials=[{"bsdf": "diffuse", "params": {"transmission": [2, 2, 2]}}],
                    model=list(base["model"]) + [occ])
    opaque_occ = dict(base, materials=[{"bsdf": "diffuse"}], model=list(base["model"]) + [occ])
    render = oracle.render if integrator == "wavefront" else oracle.render_megakernel
    imgs = []
    for cfg in (base, with_occ, opaque_occ):
        s = krr.Scene(cfg)
        imgs.append(render(s.desc, s.camera(0), 48, 48, 2, spp=2, max_depth=4)[0])
    assert np.array_equal(imgs[0], imgs[1])
    assert imgs[2][..., :3].mean() < 0.8 * imgs[0][..., :3].mean()  # the same quad, opaque, shades the room
