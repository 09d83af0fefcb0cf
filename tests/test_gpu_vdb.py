"""GPU parity of an OpenVDB volume (OpenVDBImporter::import,
src/scene/openvdb.cpp:8-67): a scattering density grid read from a .vdb file
(written by tests/vdb_writer.py: half-float leaves, an active tile, a
scale-translate map), placed by its node transform inside its null-material
box, lit by a quad emitter above a diffuse floor, with the camera outside
the box.  Both the persistent per-path launch and the wavefront stages must
take the same decisions as the CPU checker: equal ray counts and every pixel
within 1e-4 relative.
"""
import numpy as np
import pytest

from test_gpu_render import assert_images_match, gpu_render
from vdb_writer import write_vdb

pytestmark = pytest.mark.gpu


def _volume(tmp_path):
    rng = np.random.default_rng(12)
    leaves = {}
    for x in (0, 8, 16):
        for y in (0, 8):
            for z in (0, 8, 16):
                vals = (np.round(rng.uniform(0, 3, 512) * 64) / 64).astype(np.float32)
                act = rng.uniform(0, 1, 512) < 0.8
                leaves[(x, y, z)] = (np.where(act, vals, 0.0).astype(np.float32), act)
    g = {"name": "density", "leaves": leaves, "tiles": {((8, 16, 8), 8): (1.25, True)}, "half": True,
         "voxel": (0.05, 0.05, 0.05), "translate": (-0.6, -0.5, -0.6)}
    return write_vdb(str(tmp_path / "vol.vdb"), [g])


@pytest.mark.parametrize("wd", [0, 100])
def test_vdb_volume_parity(krr, oracle, gpu, tmp_path, wd):
    print(f"case: vdb volume 64x48 spp 4 depth 8 wavefront_depths={wd}", flush=True)
    path = _volume(tmp_path)
    s = krr.Scene({"materials": [{"bsdf": "diffuse", "diffuse": [0.6, 0.5, 0.4]}],
                   "model": [{"model": path, "translate": [0.1, 0.0, -0.2], "scale": [1.5, 1.5, 1.5],
                              "params": {"sigma_t": [4, 5, 6], "albedo": [0.9, 0.8, 0.7], "g": 0.4}},
                             {"type": "quad", "corners": [[-1, 2, -1], [1, 2, -1], [1, 2, 1], [-1, 2, 1]],
                              "material": 0, "emission": [6, 6, 6]},
                             {"type": "quad", "corners": [[-4, -0.8, 4], [4, -0.8, 4], [4, -0.8, -4], [-4, -0.8, -4]],
                              "material": 0}],
                   "resolution": [64, 48],
                   "camera": {"eye": [0, 0.4, 3.0], "target": [0, 0, 0]}})
    params = {"spp": 4, "max_depth": 8, "wavefront_depths": wd}
    g, gs = gpu_render(krr, gpu, s, 64, 48, params)
    o, os_ = oracle.render(s.desc, s.camera(0), 64, 48, 1, spp=4, max_depth=8)
    print("gpu", gs, "cpu", os_, flush=True)
    assert gs["closest_rays"] == os_["closest_rays"] and gs["shadow_rays"] == os_["shadow_rays"]
    assert_images_match(g, o)
    assert o[..., :3].max() > 0
