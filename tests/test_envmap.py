"""Image environment lights (rt::InfiniteLight with a texture,
src/core/light.h:208-259): loading, lat-long mapping and bilinear lookup.

Closed forms pin the checker's lookup.  A diffuse plane of albedo a under
an environment whose image is a ramp reflects a / pi * integral(L cos) over
its hemisphere:

* rows valued j (a ramp in v = acos(y) / pi), plane facing +y:
  a (H/4 - 1/2) + a pi^2 / (24 H); the same image turned 180 degrees about x
  (the plane now sees the other pole): a (3H/4 - 1/2) - a pi^2 / (24 H);
* columns valued i (a ramp in u = atan2(x, -z) / 2pi + 1/2), plane facing +x:
  a (3W/4 - 1/2) - a pi^2 / (24 W); facing -x: a (W/4 - 1/2) + a pi^2 / (24 W).

The main terms fix the mapping's axes, signs and the half-texel offset of
a normalised-coordinate lookup; the pi^2 / 24 terms are the wrap-around
rows and columns at the seam and the poles.  With NEE and without, the
checker's estimate must match within 5 standard errors.
"""
import math
import os

import numpy as np
import pytest

from conftest import CORNELL
from test_oracle_integrator import _check_mean

A = 0.5


def _ramp(h, w, axis):
    j, i = np.mgrid[0:h, 0:w].astype(np.float32)
    v = j if axis == "v" else i
    return np.ascontiguousarray(np.repeat(v[..., None], 3, -1))


def _plane_scene(krr, img_path, normal, rotate=None):
    big = 20.0
    if normal == "+y":
        corners = [[-big, 0, big], [big, 0, big], [big, 0, -big], [-big, 0, -big]]
        cam = {"eye": [0, 1, 0], "target": [0, 0, -0.3]}
    elif normal == "+x":
        corners = [[0, -big, big], [0, -big, -big], [0, big, -big], [0, big, big]]
        cam = {"eye": [1, 0, 0], "target": [0, 0.3, 0], "up": [0, 0, 1]}
    else:  # -x
        corners = [[0, -big, -big], [0, -big, big], [0, big, big], [0, big, -big]]
        cam = {"eye": [-1, 0, 0], "target": [0, 0.3, 0], "up": [0, 0, 1]}
    env = {"image": str(img_path), "tint": [1, 1, 1], "scale": 1.0}
    if rotate is not None:
        env["rotate"] = rotate
    return krr.Scene({"materials": [{"bsdf": "diffuse", "diffuse": [A] * 3}],
                      "model": [{"type": "quad", "corners": corners, "material": 0}],
                      "environment": env, "resolution": [32, 32], "camera": cam})


CASES = [  # (ramp axis, image h, w, plane normal, rotate, expected)
    ("v", 64, 8, "+y", None, A * (64 / 4 - 0.5) + A * math.pi ** 2 / (24 * 64)),
    ("v", 64, 8, "+y", [0, 1, 0, 0], A * (3 * 64 / 4 - 0.5) - A * math.pi ** 2 / (24 * 64)),
    ("u", 8, 64, "+x", None, A * (3 * 64 / 4 - 0.5) - A * math.pi ** 2 / (24 * 64)),
    ("u", 8, 64, "-x", None, A * (64 / 4 - 0.5) + A * math.pi ** 2 / (24 * 64)),
]


@pytest.mark.parametrize("axis,h,w,normal,rotate,expected", CASES)
@pytest.mark.parametrize("nee", [1, 0])
def test_ramp_environment_closed_form(oracle, krr, tmp_path, axis, h, w, normal, rotate, expected, nee):
    path = tmp_path / "ramp.npy"
    np.save(path, _ramp(h, w, axis))
    s = _plane_scene(krr, path, normal, rotate)
    cam = s.camera(0)
    img0, _ = oracle.render(s.desc, cam, 32, 32, 1, spp=1, max_depth=0)
    assert np.all(img0[..., :3] == 0)  # every camera ray hits the plane
    img, _ = oracle.render(s.desc, cam, 32, 32, 1, spp=64, max_depth=1, nee=nee, rr=1.0, precision=64)
    _check_mean(img, expected, f"ramp {axis} {normal} rotate={rotate} nee={nee}")
    mk, _ = oracle.render_megakernel(s.desc, cam, 32, 32, 1, spp=64, max_depth=1, nee=nee, rr=1.0, precision=64)
    _check_mean(mk, expected, f"megakernel ramp {axis} {normal} rotate={rotate} nee={nee}")


def _write_pfm(path, rgb, little=True):
    h, w, _ = rgb.shape
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n{-1.0 if little else 1.0}\n".encode())
        data = rgb[::-1].astype("<f4" if little else ">f4")  # PFM rows run bottom-up
        f.write(data.tobytes())


def _light_image(krr, s):
    d = s.desc
    assert d.n_infinite_lights == 1
    il = d.infinite_lights[0]
    n = il.image_width * il.image_height * 4
    return np.ctypeslib.as_array(il.image_rgba, shape=(n,)).reshape(il.image_height, il.image_width, 4).copy(), il


@pytest.mark.parametrize("fmt", ["npy3", "npy4", "npy1", "pfm_le", "pfm_be"])
def test_environment_image_files(krr, tmp_path, fmt):
    """Every accepted format reads to RGBA rows from the top of the image
    (alpha 1 where the file has none; grey replicated)."""
    rng = np.random.default_rng(3)
    rgba = rng.random((5, 7, 4), dtype=np.float32)
    if fmt == "npy4":
        p = tmp_path / "e.npy"
        np.save(p, rgba)
        want = rgba
    elif fmt == "npy3":
        p = tmp_path / "e.npy"
        np.save(p, rgba[..., :3].copy())
        want = np.concatenate([rgba[..., :3], np.ones((5, 7, 1), np.float32)], -1)
    elif fmt == "npy1":
        p = tmp_path / "e.npy"
        np.save(p, rgba[..., 0].copy())
        want = np.concatenate([np.repeat(rgba[..., :1], 3, -1), np.ones((5, 7, 1), np.float32)], -1)
    else:
        p = tmp_path / "e.pfm"
        _write_pfm(p, rgba[..., :3], little=fmt == "pfm_le")
        want = np.concatenate([rgba[..., :3], np.ones((5, 7, 1), np.float32)], -1)
    quad = {"type": "quad", "corners": [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], "material": 0}
    s = krr.Scene({"model": [quad], "materials": [{"bsdf": "diffuse"}], "environment": {"image": str(p), "scale": 2.0}})
    img, il = _light_image(krr, s)
    assert np.array_equal(img, want)
    assert il.scale == 2.0 and list(il.rotation) == [1, 0, 0, 0, 1, 0, 0, 0, 1]


def test_environment_forms(krr, tmp_path):
    """The reference's forms: "environment": "<image>" (krrscene.cpp:268-274)
    and a light node {"type": "infinite", "texture", "color", "scale"} whose
    node rotation turns the image; paths relative to "base_dir"."""
    np.save(tmp_path / "sky.npy", np.ones((4, 8, 3), np.float32))
    quad = {"type": "quad", "corners": [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], "material": 0}
    mat = [{"bsdf": "diffuse"}]
    s = krr.Scene({"model": [quad], "materials": mat, "environment": "sky.npy", "base_dir": str(tmp_path)})
    img, il = _light_image(krr, s)
    assert img.shape == (4, 8, 4) and il.scale == 1.0 and list(il.tint) == [1, 1, 1]
    h = math.sqrt(0.5)
    s = krr.Scene({"model": [quad, {"type": "light", "rotate": [h, 0, h, 0],
                                    "params": {"type": "infinite", "texture": "sky.npy", "color": [1, 2, 3], "scale": 4}}],
                   "materials": mat, "base_dir": str(tmp_path)})
    img, il = _light_image(krr, s)
    R = np.array(il.rotation).reshape(3, 3)
    assert np.allclose(R, [[0, 0, 1], [0, 1, 0], [-1, 0, 0]], atol=1e-6)  # 90 degrees about +y
    assert list(il.tint) == [1, 2, 3] and il.scale == 4
    for bad in ["sky.exr", "missing.npy"]:
        with pytest.raises(krr.KrrError):
            krr.Scene({"model": [quad], "materials": mat, "environment": bad, "base_dir": str(tmp_path)})
    np.save(tmp_path / "f64.npy", np.ones((4, 8, 3), np.float64))
    with pytest.raises(krr.KrrError):
        krr.Scene({"model": [quad], "materials": mat, "environment": "f64.npy", "base_dir": str(tmp_path)})
