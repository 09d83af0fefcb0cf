"""The C++ render-pass layer (kiraray_amd/host/): RenderPass factory and
registration, WavefrontPathTracer params, RenderApp config loading, and one
render through the host loop checked against the CPU checker."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import CORNELL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "kiraray_amd", "lib", "host_tests")


def _run(*args, timeout=120):
    if not os.path.exists(EXE):
        pytest.fail("kiraray_amd/lib/host_tests is not built (run __graft_entry__.build())")
    r = subprocess.run([EXE, *map(str, args)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=timeout)
    print(r.stderr, flush=True)
    return r


@pytest.mark.parametrize("case", ["factory", "params", "config"])
def test_host_layer(case):
    r = _run(case)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_render_app_matches_checker(oracle, krr, tmp_path):
    """Three frames of the Cornell box through RenderApp -> WavefrontPathTracer
    (1-based frame index, so the last frame is frame 3) equal the checker's
    frame 3."""
    print("case: RenderApp cornell 80x60 frames 1..3 spp 2", flush=True)
    out = tmp_path / "film.bin"
    r = _run("render", 80, 60, 3, out)
    assert r.returncode == 0, r.stderr
    st = json.loads(r.stdout.strip().splitlines()[-1])
    assert st["frame_index"] == 3
    g = np.fromfile(out, np.float32).reshape(60, 80, 4)
    s = krr.Scene(CORNELL)
    cam = s.camera(0)
    cam.aspect_ratio = 80 / 60  # RenderApp::resize sets width / height
    o, ost = oracle.render(s.desc, cam, 80, 60, 3, spp=2, max_depth=8)
    assert st["closest_rays"] == ost["closest_rays"] and st["shadow_rays"] == ost["shadow_rays"]
    d = np.abs(g[..., :3] - o[..., :3])
    assert (d <= 1e-4 * np.abs(o[..., :3]) + 1e-6).all(), d.max()
    assert np.all(g[..., 3] == 1.0)


@pytest.mark.gpu
def test_accumulate_and_error_measure_passes(oracle, krr, tmp_path):
    """[WavefrontPathTracer, AccumulatePass (double), ErrorMeasurePass (relMSE
    against the 65,536-spp FP64 Cornell image every 4 frames, saved)] through
    RenderApp for 32 frames: the saved series has timesteps 4, 8, .., 32; the
    accumulated image equals the checker's frames 1..32 averaged in double;
    the last value equals the relMSE of that image; the error falls ~1/frames."""
    print("case: RenderApp cornell 256 + AccumulatePass + ErrorMeasurePass, 32 frames", flush=True)
    from test_relmse import reference_rgba, relmse_np
    ref = reference_rgba()
    ref_path = tmp_path / "ref.npy"
    np.save(ref_path, np.ascontiguousarray(ref[..., :3], np.float32))
    out = tmp_path / "error" / "series.json"
    img_path = tmp_path / "accum.bin"
    r = _run("measure", 32, ref_path, out, img_path, timeout=300)
    assert r.returncode == 0, r.stderr
    series = json.loads(out.read_text())
    assert series["timesteps"] == [4, 8, 12, 16, 20, 24, 28, 32]
    assert len(series["timepoints"]) == 8 and all(b >= a for a, b in zip(series["timepoints"], series["timepoints"][1:]))
    vals = [d["RelMSE"] for d in series["data"]]
    g = np.fromfile(img_path, np.float32).reshape(256, 256, 4)
    s = krr.Scene(CORNELL)
    acc = np.zeros((256, 256, 3), np.float64)
    for f in range(1, 33):
        o, _ = oracle.render(s.desc, s.camera(0), 256, 256, f, spp=1, max_depth=8)
        acc += o[..., :3].astype(np.float64)
    mean = (acc / 32).astype(np.float32)
    assert np.array_equal(g[..., :3], mean)
    rel = relmse_np(g, ref)
    print("series", [round(v, 4) for v in vals], "host", rel)
    assert abs(vals[-1] - rel) <= 1e-6 * rel
    assert 4.0 < vals[0] / vals[-1] < 12.0  # 4 -> 32 frames: ~8x


@pytest.mark.gpu
def test_render_app_megakernel_matches_checker(oracle, krr, tmp_path):
    """Three frames of the Cornell box through RenderApp ->
    MegakernelPathTracer: the colour texture holds frame 3, equal to the
    checker's megakernel frame 3."""
    print("case: RenderApp cornell 80x60 MegakernelPathTracer frames 1..3 spp 2", flush=True)
    out = tmp_path / "film_mk.bin"
    r = _run("render_mk", 80, 60, 3, out)
    assert r.returncode == 0, r.stderr
    st = json.loads(r.stdout.strip().splitlines()[-1])
    g = np.fromfile(out, np.float32).reshape(60, 80, 4)
    s = krr.Scene(CORNELL)
    cam = s.camera(0)
    cam.aspect_ratio = 80 / 60
    o, ost = oracle.render_megakernel(s.desc, cam, 80, 60, 3, spp=2, max_depth=8)
    assert st["closest_rays"] == ost["closest_rays"] and st["shadow_rays"] == ost["shadow_rays"]
    d = np.abs(g[..., :3] - o[..., :3])
    assert (d <= 1e-4 * np.abs(o[..., :3]) + 1e-6).all(), d.max()
