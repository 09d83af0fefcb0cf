"""Pins of the checker's megakernel path tracer (oracle_render_megakernel,
src/render/megakernel/device.cu:50-195).

It shares the sampling and shading routines pinned in test_oracle_pins.py
with the wavefront checker; these tests fix how it combines them, through
quantities its estimator must reproduce whatever its random choices:

* Emissive furnace (closed cube, every wall emits Le and reflects with
  albedo a): radiance Le * sum_{k=0..D} a^k for max depth D, with NEE + MIS
  (power heuristic), BSDF sampling alone, and any roulette probability.
* Sky over a ground plane: a diffuse plane of albedo a under a constant
  infinite light Le reflects a * Le; this goes through the miss-side MIS
  weight and the infinite light's light samples.
* The Cornell box against the 65,536-spp FP64 image of the wavefront
  integrator (tests/golden/): both integrators estimate the same
  depth-limited path integral, so block means must agree within their
  statistical error, with NEE on and off.

A wrong MIS weight, light-selection or light pdf, or roulette compensation
biases at least one of these.  Tolerances are 5 standard errors.
"""
import numpy as np
import pytest

from conftest import CORNELL
from test_oracle_integrator import _check_mean, _furnace_scene
from test_relmse import reference_rgba


@pytest.mark.parametrize("nee,rr,depth,precision", [
    (1, 0.8, 5, 32),   # reference defaults: NEE + MIS, roulette from the first vertex
    (0, 0.8, 5, 32),   # BSDF sampling only
    (1, 1.0, 4, 32),   # no roulette (no draw is taken for it)
    (1, 0.5, 7, 32),   # heavy roulette
    (1, 0.8, 3, 64),   # FP64
    (1, 0.8, 0, 32),   # emission only
])
def test_megakernel_furnace(oracle, krr, nee, rr, depth, precision):
    albedo, le = 0.5, 1.0
    s = _furnace_scene(krr, albedo, le)
    img, st = oracle.render_megakernel(s.desc, s.camera(0), 48, 48, 1, spp=32, max_depth=depth, nee=nee, rr=rr,
                                       precision=precision)
    expected = le * sum(albedo ** k for k in range(depth + 1))
    _check_mean(img, expected, f"megakernel furnace nee={nee} rr={rr} D={depth} fp{precision}")
    if depth == 0:
        assert np.all(img[..., :3] == np.float32(le))
    if nee and depth > 0:
        assert st["shadow_rays"] > 0
    elif not nee:
        assert st["shadow_rays"] == 0


def test_megakernel_furnace_sees_albedo(oracle, krr):
    s = _furnace_scene(krr, 0.8, 2.0)
    img, _ = oracle.render_megakernel(s.desc, s.camera(0), 48, 48, 1, spp=48, max_depth=6)
    _check_mean(img, 2.0 * sum(0.8 ** k for k in range(7)), "megakernel furnace a=0.8 Le=2 D=6")


def _sky_scene(krr, albedo, le):
    # Upward rays from the plane see only sky whatever its size; 20 keeps its
    # triangles small enough that the fixed 1e-4 ray offset clears them (with
    # 1e4-sized triangles float rounding in the watertight test lets some
    # reflected rays hit the plane again).
    big = 20.0
    return krr.Scene({"materials": [{"bsdf": "diffuse", "diffuse": [albedo] * 3}],
                      "model": [{"type": "quad", "corners": [[-big, 0, big], [big, 0, big], [big, 0, -big], [-big, 0, -big]],
                                 "material": 0}],
                      "environment": {"tint": [le] * 3, "scale": 1.0},
                      "resolution": [48, 48],
                      "camera": {"eye": [0, 1, 0], "target": [0, 0, -0.3]}})


@pytest.mark.parametrize("nee", [1, 0])
def test_megakernel_sky_over_plane(oracle, krr, nee):
    """Every camera ray of this view hits the plane (checked on the depth-0
    image); the plane then shows a * Le."""
    albedo, le = 0.6, 1.5
    s = _sky_scene(krr, albedo, le)
    cam = s.camera(0)
    img0, _ = oracle.render_megakernel(s.desc, cam, 48, 48, 1, spp=1, max_depth=0)
    assert np.all(img0[..., :3] == 0)  # the ground does not emit: nothing escapes at depth 0
    img, _ = oracle.render_megakernel(s.desc, cam, 48, 48, 1, spp=32, max_depth=3, nee=nee)
    _check_mean(img, albedo * le, f"megakernel sky over plane nee={nee}")


@pytest.mark.parametrize("nee,spp", [(1, 64), (0, 256)])
def test_megakernel_cornell_matches_wavefront_reference(oracle, krr, nee, spp):
    """Every other row and column of the 256 x 256 Cornell box, FP64, against
    the 65,536-spp wavefront image: the mean difference over each 64 x
    64-pixel block (32 x 32 rendered pixels) is within 5 standard errors of
    that block, and so is the whole image's."""
    s = krr.Scene(CORNELL)
    ref = reference_rgba()[..., :3].astype(np.float64)
    img, st = oracle.render_megakernel(s.desc, s.camera(0), 256, 256, 3, spp=spp, max_depth=8, nee=nee, rr=0.8,
                                       precision=64, row_step=2, pixel_step=2)
    print(f"megakernel cornell nee={nee} spp={spp}: {st['seconds']:.1f} s", flush=True)
    rows = 255 - np.arange(0, 256, 2)  # rendered rows y = 0, 2, ..; film row H-1-y
    d = (img[rows][:, ::2, :3].astype(np.float64) - ref[rows][:, ::2]).reshape(4, 32, 4, 32, 3)
    blk = d.mean(axis=(1, 3))
    se = d.std(axis=(1, 3)) / np.sqrt(32 * 32)
    print(f"  block bias / se: max {np.max(np.abs(blk) / (se + 1e-12)):.2f}", flush=True)
    assert np.all(np.abs(blk) <= 5 * se + 1e-6)
    v = d.reshape(-1)
    assert abs(v.mean()) <= 5 * v.std() / np.sqrt(v.size)
