"""relMSE against the 65,536-spp FP64 Cornell reference
(tests/golden/cornell_256_fp64_65536spp.npz, made by make_cornell_reference.py
with the CPU checker): the reference metric (src/render/passes/errormeasure/
metrics.cu:80-90) falls as 1/spp for the unbiased FP32 tracer, and the GPU
metric kernel equals the host computation."""
import os

import numpy as np
import pytest

from conftest import CORNELL

REF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cornell_256_fp64_65536spp.npz")


def reference_rgba():
    z = np.load(REF)
    img = z["image"]
    assert img.shape == (256, 256, 3) and int(z["spp"]) == 65536
    return np.concatenate([img, np.ones((256, 256, 1), np.float32)], -1)


def relmse_np(img, ref):
    """metrics.cu:80-90: mean over RGB of ((y - r) / r)^2 (0 where r == 0),
    clamped at 100 per pixel, averaged over pixels, in float64."""
    y = img[..., :3].astype(np.float64)
    r = ref[..., :3].astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        e = np.where(r != 0, ((y - r) / r) ** 2, 0.0)
    return float(np.minimum(e.mean(-1), 100.0).mean())


def test_reference_image_is_plausible():
    ref = reference_rgba()
    assert np.isfinite(ref).all() and (ref[..., :3] >= 0).all()
    assert 0.15 < ref[..., :3].mean() < 0.35
    # left wall red, right wall green (seen from the front, film rows flipped)
    h = ref[128 - 20:128 + 20]
    assert h[:, 5:20, 0].mean() > 2 * h[:, 5:20, 1].mean()
    assert h[:, 236:251, 1].mean() > 1.5 * h[:, 236:251, 0].mean()


def test_relmse_falls_as_one_over_spp(oracle, krr):
    """Unbiased estimator: relMSE(spp) ~ c / spp.  4 -> 16 spp must cut it
    by 4 within noise (factor 2.8 .. 5.5)."""
    s = krr.Scene(CORNELL)
    ref = reference_rgba()
    r = {}
    for spp in (4, 16):
        img, _ = oracle.render(s.desc, s.camera(0), 256, 256, 1, spp=spp, max_depth=8)
        r[spp] = oracle.error_metric(3, img, ref)
        assert abs(r[spp] - relmse_np(img, ref)) <= 1e-9 + 1e-6 * r[spp]
    print(f"relMSE 4 spp {r[4]:.5f}, 16 spp {r[16]:.5f}, ratio {r[4] / r[16]:.2f}")
    assert 2.8 < r[4] / r[16] < 5.5


@pytest.mark.gpu
def test_gpu_relmse_metric(krr, oracle, gpu):
    """16-spp GPU image: the device relMSE kernel equals the host formula,
    and (the images being identical) the checker's own relMSE."""
    from test_gpu_render import gpu_render
    print("case: relMSE cornell 256 16 spp", flush=True)
    s = krr.Scene(CORNELL)
    g, _ = gpu_render(krr, gpu, s, 256, 256, {"spp": 16, "max_depth": 8})
    ref = reference_rgba()
    gb, rb = gpu.DeviceBuffer.from_numpy(g), gpu.DeviceBuffer.from_numpy(ref)
    dev = krr.error_metric(3, gb.ptr, rb.ptr, 256 * 256)
    host = relmse_np(g, ref)
    o, _ = oracle.render(s.desc, s.camera(0), 256, 256, 1, spp=16, max_depth=8)
    print(f"relMSE device {dev:.6f} host {host:.6f} checker image {relmse_np(o, ref):.6f}")
    assert abs(dev - host) <= 1e-6 * host
    assert abs(host - relmse_np(o, ref)) <= 1e-4 * host
    assert host < 0.5


@pytest.mark.gpu
@pytest.mark.parametrize("mode,precision,max_count", [(0, 32, 0), (0, 64, 0), (1, 32, 0), (0, 32, 3)])
def test_gpu_accumulate_pass(krr, gpu, mode, precision, max_count):
    """krr_accumulate against the AccumulatePass arithmetic (accumulate.cu:30-79)
    over 5 random frames."""
    print(f"case: accumulate mode={mode} precision={precision} max={max_count}", flush=True)
    rng = np.random.default_rng(mode * 10 + precision + max_count)
    n = 1000
    acc = gpu.DeviceBuffer(n * 4 * (4 if precision == 32 else 8))
    dt = np.float32 if precision == 32 else np.float64
    a = np.zeros((n, 4), dt)
    for count in range(5):
        f = rng.random((n, 4)).astype(np.float32)
        fb = gpu.DeviceBuffer.from_numpy(f)
        krr.accumulate(acc.ptr, fb.ptr, n, count, max_count, mode, precision)
        gpu.synchronize()
        out = fb.to_numpy(np.float32, (n, 4))
        w = dt(1) / dt(count + 1)
        c = f.astype(dt)
        if count == 0:
            a = c.copy()
        elif mode == 1:
            a = a + (c - a) * w
        elif not max_count or count < max_count:
            a = a + c
        exp = a if mode == 1 else a * w
        np.testing.assert_allclose(out, exp.astype(np.float32), rtol=2e-6, atol=1e-7)
