"""Ablation of the radiance cache on the 1.31 M-triangle field at 1920x1080:
what the trained cache gains over the same tracer without it, at equal time.

Error reference: the FP64 checker image of the same view at 2,048 spp
(tests/golden/icosphere_field_1920x1080_fp64_2048spp.npz, made by
tests/golden/make_icosphere_reference.py --full).  The plain tracer renders
frames 1.., the cached tracer trains online for 48 frames and then renders
frames with training continuing; each frame is timed on the host (begin_frame
+ render + synchronize, training included).  With the cached tracer's frames
taken up to the wall time the plain tracer spent, the relMSE of its running
average (the reference's metric, metrics.cu:80-90) must be lower than the
plain tracer's.  The warm-up frames are not on the clock: the cache is
compared in its steady state, as in interactive use.
"""
import os
import time

import numpy as np
import pytest

from conftest import ROOT, icosphere_scene

pytestmark = pytest.mark.gpu

REFERENCE = os.path.join(ROOT, "tests", "golden", "icosphere_field_1920x1080_fp64_2048spp.npz")
W, H = 1920, 1080


def _running_relmse(krr, gpu, scene, params, first, frames, ref, warm=0, budget_s=None):
    """(frame times, relMSE of the running average after each frame)."""
    n = W * H
    film = gpu.DeviceBuffer(n * 16)
    acc = gpu.DeviceBuffer(n * 32)
    wf = krr.WavefrontPathTracer(params)
    wf.set_scene(scene)
    wf.resize(W, H)
    cam = scene.camera(0)
    for i in range(warm):
        wf.begin_frame(cam, first - warm + i)
        wf.render(film.ptr, 0)
    gpu.synchronize()
    times, rel = [], []
    for i in range(frames):
        t0 = time.perf_counter()
        wf.begin_frame(cam, first + i)
        wf.render(film.ptr, 0)
        gpu.synchronize()
        times.append(time.perf_counter() - t0)
        krr.accumulate(acc.ptr, film.ptr, n, i, 0, 0, 64)
        gpu.synchronize()
        rel.append(krr.error_metric(3, film.ptr, ref.ptr, n))
        if budget_s is not None and sum(times) >= budget_s:
            break
    wf.close()
    return times, rel


def test_cache_beats_plain_at_equal_time(krr, gpu):
    print("case: nrc ablation, icosphere field 1920x1080 vs FP64 2048-spp reference", flush=True)
    with np.load(REFERENCE, allow_pickle=False) as z:
        img = z["image"].astype(np.float32)
    assert img.shape == (H, W, 3)
    rgba = np.ones((H, W, 4), np.float32)
    rgba[..., :3] = img
    ref = gpu.DeviceBuffer.from_numpy(rgba)
    s = krr.Scene(icosphere_scene(res=(W, H)))
    base = {"spp": 1, "max_depth": 16, "rr": 0.8, "rr_min_depth": 3}
    tp, rp = _running_relmse(krr, gpu, s, base, 1, 12, ref)
    nrc = dict(base, nrc={"train_mod": 16, "batch": 65536, "steps_per_frame": 4, "lr": 1e-3})
    tn, rn = _running_relmse(krr, gpu, s, nrc, 49, 64, ref, warm=48, budget_s=sum(tp))
    # the cache's frames that fit in the plain tracer's wall time
    k = int(np.searchsorted(np.cumsum(tn), sum(tp), side="right"))
    print(f"plain: {len(tp)} frames in {1e3 * sum(tp):.1f} ms, relMSE {rp[-1]:.4f}; "
          f"cache: {k} frames in {1e3 * float(np.sum(tn[:k])):.1f} ms, relMSE {rn[k - 1]:.4f}", flush=True)
    assert k >= 6
    assert rn[k - 1] < rp[-1]
