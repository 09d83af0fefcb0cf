"""Closed-form pins of the CPU checker's integrator as a whole.

The component pins (test_oracle_pins.py) fix each sampling and shading
routine; these fix how the integrator combines them.  A dropped MIS weight,
a wrong light pdf, a missing Russian-roulette compensation or a wrong
transmittance would each bias an image away from the closed forms below.

* Emissive furnace: inside a closed cube whose walls all emit Le (one-sided,
  towards the inside) and reflect diffusely with albedo a, the radiance is
  the same everywhere and equals Le * sum_{k=0..D} a^k for max depth D: each
  bounce adds one more term, whatever the sampling (NEE + MIS, BSDF only,
  Russian roulette).  Hits at depths 0..D add emission and NEE from depths
  0..D-1 adds the same terms (integrator.cpp:232-255).
* Homogeneous absorber: a camera inside a medium with albedo 0 sees an
  emitter at distance t with radiance Le * exp(-sigma_t * t) (delta tracking
  absorbs with probability 1 - exp(-sigma_t t), medium.cpp:13-103).

Tolerances are statistical: 5 standard errors of the image mean, computed
from the per-pixel spread of the same image.
"""
import numpy as np
import pytest


def _cube_inward(half=1.0):
    """Closed cube [-half, half]^3, 6 faces x 2 triangles, normals inward."""
    pos, nrm, idx = [], [], []
    faces = [((1, 0, 0), (0, 1, 0), (0, 0, 1)), ((-1, 0, 0), (0, 0, 1), (0, 1, 0)),
             ((0, 1, 0), (0, 0, 1), (1, 0, 0)), ((0, -1, 0), (1, 0, 0), (0, 0, 1)),
             ((0, 0, 1), (1, 0, 0), (0, 1, 0)), ((0, 0, -1), (0, 1, 0), (1, 0, 0))]
    for n_in, u, v in faces:
        n_in, u, v = np.array(n_in, float), np.array(u, float), np.array(v, float)
        c = -n_in * half  # the face whose inward normal is n_in lies at -n_in * half
        corners = [c - u * half - v * half, c + u * half - v * half, c + u * half + v * half, c - u * half + v * half]
        base = len(pos)
        # wind so that cross(p1 - p0, p2 - p0) points along the inward normal
        if np.dot(np.cross(corners[1] - corners[0], corners[2] - corners[0]), n_in) < 0:
            corners = corners[::-1]
        pos += [list(map(float, p)) for p in corners]
        nrm += [list(map(float, n_in))] * 4
        idx += [base, base + 1, base + 2, base, base + 2, base + 3]
    return [x for p in pos for x in p], [x for n in nrm for x in n], idx


def _furnace_scene(krr, albedo, le, res=48):
    pos, nrm, idx = _cube_inward()
    return krr.Scene({"materials": [{"bsdf": "diffuse", "diffuse": [albedo] * 3}],
                      "model": [{"type": "mesh", "positions": pos, "indices": idx, "normals": nrm,
                                 "material": 0, "emission": [le] * 3}],
                      "resolution": [res, res],
                      "camera": {"eye": [0.1, -0.2, 0.3], "target": [0.4, 0.1, -1.0]}})


def _check_mean(img, expected, what):
    v = img[..., :3].astype(np.float64).ravel()
    se = v.std() / np.sqrt(v.size)
    print(f"{what}: mean {v.mean():.6f} expected {expected:.6f} (5 se = {5 * se:.2e})", flush=True)
    assert abs(v.mean() - expected) <= 5 * se + 1e-6 * expected, (v.mean(), expected, se)


@pytest.mark.parametrize("nee,rr,rr_min,depth,precision", [
    (1, 0.8, 0, 5, 32),   # reference defaults, NEE + MIS
    (0, 0.8, 0, 5, 32),   # BSDF sampling only
    (1, 0.5, 2, 7, 32),   # heavier roulette from depth 2
    (1, 0.8, 0, 3, 64),   # reference-image (FP64) mode
    (1, 0.8, 0, 0, 32),   # emission only
])
def test_emissive_furnace(oracle, krr, nee, rr, rr_min, depth, precision):
    albedo, le = 0.5, 1.0
    s = _furnace_scene(krr, albedo, le)
    img, st = oracle.render(s.desc, s.camera(0), 48, 48, 1, spp=24, max_depth=depth, nee=nee, rr=rr,
                            rr_min_depth=rr_min, precision=precision)
    expected = le * sum(albedo ** k for k in range(depth + 1))
    _check_mean(img, expected, f"furnace nee={nee} rr={rr}@{rr_min} D={depth} fp{precision}")
    if depth == 0:
        assert np.all(img[..., :3] == np.float32(le))  # every camera ray sees the inside of a wall


def test_emissive_furnace_sees_albedo(oracle, krr):
    """The closed form must also react to the albedo (a broken BSDF that
    returned the wrong reflectance would fail here even if it were unbiased
    for a = 0.5)."""
    s = _furnace_scene(krr, 0.8, 2.0)
    img, _ = oracle.render(s.desc, s.camera(0), 48, 48, 1, spp=32, max_depth=6)
    _check_mean(img, 2.0 * sum(0.8 ** k for k in range(7)), "furnace a=0.8 Le=2 D=6")


def test_homogeneous_absorber(oracle, krr):
    """Emitter quad at z = -2 seen through a purely absorbing medium
    (sigma_t = 0.7, albedo 0): pixel = Le * exp(-0.7 t), t = distance to the
    quad along the pixel's ray."""
    sigma, le, w, h = 0.7, 3.0, 40, 40
    s = krr.Scene({"materials": [{"bsdf": "diffuse", "diffuse": [0.0, 0.0, 0.0]}],
                   "media": [{"type": "homogeneous", "sigma_t": [sigma] * 3, "albedo": [0.0] * 3, "g": 0.0}],
                   "model": [{"type": "quad", "corners": [[-4, -4, -2], [4, -4, -2], [4, 4, -2], [-4, 4, -2]],
                              "material": 0, "emission": [le] * 3}],
                   "resolution": [w, h],
                   "camera": {"eye": [0, 0, 0], "target": [0, 0, -1], "medium": 0}})
    cam = s.camera(0)
    img, _ = oracle.render(s.desc, cam, w, h, 1, spp=64, max_depth=2)
    # expected radiance per pixel from the checker's own camera ray (pinned in
    # test_oracle_pins.py), at the centre of the pixel's sample footprint
    exp_img = np.zeros((h, w))
    for y in range(h):
        for x in range(w):
            o, d = oracle.camera_ray(cam, x, y, w, h, (0.5, 0.5))
            d = np.asarray(d, np.float64)
            t = (-2.0 - o[2]) / d[2]
            exp_img[h - 1 - y, x] = le * np.exp(-sigma * t * np.linalg.norm(d))
    v = img[..., 0].astype(np.float64)
    # per-pixel values are Le or 0 (Bernoulli): compare the image means
    var = (exp_img * (le - exp_img)).sum() / (64 * v.size ** 2)
    print(f"absorber: mean {v.mean():.5f} expected {exp_img.mean():.5f} (5 se {5 * np.sqrt(var):.2e})")
    assert abs(v.mean() - exp_img.mean()) <= 5 * np.sqrt(var) + 2e-3 * exp_img.mean()
    assert set(np.unique(img[..., 0] * 64).round(3)) <= set(np.arange(65) * le)  # k of 64 samples reach the light
