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
