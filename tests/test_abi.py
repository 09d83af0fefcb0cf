"""C ABI checks that need no GPU: the library loads, exports every function
include/*.h declares, parses configs, builds scenes and reports errors as
codes (the reference logs and exits instead, src/util/check.h:8-42)."""
import ctypes as C
import glob
import json
import os
import re

import pytest

from conftest import CORNELL, ROOT


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(krr_[a-z_0-9]+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_exports_every_declared_function(krr):
    lib = C.CDLL(os.path.join(ROOT, "kiraray_amd", "lib", "libkrr_hip.so"))
    names = _declared_functions()
    assert len(names) >= 24
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a prototype for each of them
    from kiraray_amd._abi import PROTOTYPES
    assert names <= set(PROTOTYPES), names - set(PROTOTYPES)


def test_version_and_device_count(krr):
    lib = krr.load_library()
    assert lib.krr_abi_version() == 2
    assert lib.krr_device_count() >= 0


def test_pass_params_roundtrip(krr):
    wf = krr.WavefrontPathTracer({"nee": False, "max_depth": 16, "rr": 0.5, "spp": 3, "rr_min_depth": 3})
    p = wf.to_json()
    assert p["nee"] is False and p["max_depth"] == 16 and p["spp"] == 3 and p["rr_min_depth"] == 3
    assert abs(p["rr"] - 0.5) < 1e-7
    d = krr.WavefrontPathTracer(None).to_json()
    # reference defaults (integrator.h:96-103)
    assert d["nee"] is True and d["enable_medium"] is True and d["max_depth"] == 10
    assert abs(d["rr"] - 0.8) < 1e-7 and d["enable_clamp"] is False and d["clamp_max"] == 1000
    assert d["wavefront_depths"] == 0 and d["stack_lds_entries"] == 16 and d["leaf_size"] == 4
    assert d["occupancy"] == 1 and d["graph"] is True and d["bvh_builder"] == "device"
    assert d["refill_lanes"] == -1 and d["leaf_share"] == 35 and d["lds_nodes"] == 0


@pytest.mark.parametrize("bad", ["{", "[1,2]", '{"max_depth": "ten"}', '{"spp": 0}', '{"block_size": 100}',
                                 '{"stack_lds_entries": -1}', '{"stack_lds_entries": 97}', '{"wavefront_depths": -1}',
                                 '{"occupancy": 0}', '{"occupancy": 1.5}', '{"bvh_builder": "gpu"}',
                                 '{"nrc": {"spread": -1}}', '{"refill_lanes": -2}', '{"refill_lanes": 65}',
                                 '{"lds_nodes": -1}', '{"lds_nodes": 1025}',
                                 '{"leaf_share": -1}', '{"leaf_share": 101}'])
def test_bad_params_are_reported(krr, bad):
    with pytest.raises(krr.KrrError) as e:
        krr.WavefrontPathTracer(bad)
    assert e.value.code == -2


def test_render_without_scene_is_an_error(krr):
    wf = krr.WavefrontPathTracer(None)
    with pytest.raises(krr.KrrError) as e:
        wf.render(0)
    assert e.value.code == -3


def test_cornell_scene_layout(krr):
    s = krr.Scene(CORNELL)
    assert s.n_triangles == 36
    d = s.desc
    emissive = [i for i in range(d.n_meshes) if any(d.meshes[i].emission[c] > 0 for c in range(3))]
    assert len(emissive) == 1 and d.meshes[emissive[0]].n_triangles == 2
    assert d.meshes[emissive[0]].has_normals == 1
    cam = s.camera(0)
    assert abs(cam.aspect_ratio - 1.0) < 1e-7 and abs(cam.film_size[1] - 24.0) < 1e-6
    assert s.view_size(0) == (256, 256)


def test_icosphere_field_counts(krr):
    s = krr.Scene({"model": [{"type": "icosphere_field", "count": 10, "level": 2, "lights": 4}]})
    assert s.n_triangles == 10 * 320 + 2 + 8
    mats = [s.desc.materials[i].type for i in range(s.desc.n_materials)]
    assert set(mats) <= {1, 2, 3}


def test_volume_and_views(krr):
    s = krr.Scene({"model": [{"type": "icosphere_field", "count": 4, "level": 1}],
                   "media": [{"type": "value_noise_grid", "res": 16}],
                   "camera": {"medium": 0},
                   "views": {"count": 6, "sizes": [16, 32]}})
    assert s.n_views == 6
    assert all(s.view_size(v)[0] in (16, 32) for v in range(6))
    m = s.desc.media[0]
    assert m.type == 1 and list(m.res) == [16, 16, 16]
    import numpy as np
    dens = np.ctypeslib.as_array(m.density, shape=(16 ** 3,))
    assert dens.max() == 1.0 and dens.min() >= 0.0
    assert s.camera(3).medium == 0


@pytest.mark.parametrize("cfg", [
    "not json", '{"model": [{"type": "teapot"}]}', '{"model": [{"type": "mesh", "positions": [0,0,0], "indices": [0,1,2]}]}',
    '{"camera": {"medium": 2}}',
    '{"model": [{"type": "quad", "corners": [[0,0,0],[1,0,0],[1,1,0]]}]}'])
def test_bad_scene_configs(krr, cfg):
    with pytest.raises(krr.KrrError) as e:
        krr.Scene(cfg)
    assert e.value.code == -2


def test_no_device_is_reported_not_crashed(krr):
    lib = krr.load_library()
    if lib.krr_device_count() > 0:
        pytest.skip("a HIP device is visible")
    s = krr.Scene(CORNELL)
    wf = krr.WavefrontPathTracer(None)
    with pytest.raises(krr.KrrError) as e:
        wf.set_scene(s)
    assert e.value.code == -8
    with pytest.raises(krr.KrrError) as e:
        krr.Accel(s.desc)
    assert e.value.code == -8


def test_views_argument_errors(krr):
    """set_views / render_views / autotune reject bad arguments before any
    device work."""
    import ctypes as C
    lib = krr.load_library()
    wf = krr.WavefrontPathTracer(None)
    assert lib.krr_wf_set_views(wf._h, None, 0) == krr._abi.KRR_ERR_INVALID_ARG
    v = (krr.krr_view * 1)()
    v[0].width, v[0].height, v[0].film_dev = 0, 4, 16
    assert lib.krr_wf_set_views(wf._h, v, 1) == krr._abi.KRR_ERR_INVALID_ARG
    v[0].width, v[0].film_dev = 4, None
    assert lib.krr_wf_set_views(wf._h, v, 1) == krr._abi.KRR_ERR_INVALID_ARG
    assert lib.krr_wf_render_views(wf._h, 1, None) == krr._abi.KRR_ERR_NO_SCENE
    assert lib.krr_wf_autotune(wf._h, 0, None, 0, None) == krr._abi.KRR_ERR_NO_SCENE
    wf.close()


def test_accel_builder_option_is_checked(krr):
    """An unknown "builder" fails as a bad config before any device work."""
    s = krr.Scene(CORNELL)
    with pytest.raises(krr.KrrError) as e:
        krr.Accel(s.desc, {"builder": "gpu"})
    assert e.value.code == -2


def test_megakernel_pass_params(krr):
    """MegakernelPathTracer params: the reference's keys and defaults
    (pathtracer.h:37-49), this library's additions, and their checks."""
    d = krr.MegakernelPathTracer(None).to_json()
    assert d["nee"] is True and d["max_depth"] == 10 and abs(d["rr"] - 0.8) < 1e-7
    assert d["spp"] == 1 and d["block_size"] == 256 and d["stack_lds_entries"] == 16 and d["bvh_builder"] == "device"
    p = krr.MegakernelPathTracer({"nee": False, "max_depth": 5, "rr": 1.0, "spp": 2, "leaf_size": 2}).to_json()
    assert p["nee"] is False and p["max_depth"] == 5 and p["rr"] == 1.0 and p["spp"] == 2 and p["leaf_size"] == 2
    for bad in ["{", '{"max_depth": -1}', '{"rr": 0}', '{"rr": 1.5}', '{"spp": 0}', '{"block_size": 96}',
                '{"occupancy": 0}', '{"stack_lds_entries": 97}', '{"bvh_builder": "gpu"}']:
        with pytest.raises(krr.KrrError) as e:
            krr.MegakernelPathTracer(bad)
        assert e.value.code == -2, bad
    mk = krr.MegakernelPathTracer(None)
    with pytest.raises(krr.KrrError) as e:
        mk.render(0)
    assert e.value.code == -3


def test_megakernel_rejects_medium_interfaces(krr):
    """Meshes without material are medium interfaces, which the megakernel
    integrator does not handle: set_scene refuses them before any device
    work."""
    s = krr.Scene({"model": [{"type": "cornell_box"},
                             {"type": "medium", "params": {"bound": [[-1, -1, -1], [1, 1, 1]]}}]})
    mk = krr.MegakernelPathTracer(None)
    with pytest.raises(krr.KrrError) as e:
        mk.set_scene(s)
    assert e.value.code == -7  # KRR_ERR_UNSUPPORTED


@pytest.mark.parametrize("first,count", [(-1, 2), (34, 3), (0, -1), (36, 1)])
def test_bad_mesh_range_is_rejected_before_upload(krr, first, count):
    """A mesh whose triangle range leaves the scene's triangle array is an
    invalid argument for every entry point that takes a scene, reported
    before any per-triangle table is written (and before the device check,
    so it holds without a GPU too)."""
    s = krr.Scene(CORNELL)
    d = s.desc
    meshes = (krr._abi.krr_mesh * d.n_meshes)()
    for i in range(d.n_meshes):
        meshes[i] = d.meshes[i]
    meshes[0].first_triangle, meshes[0].n_triangles = first, count
    bad = krr._abi.krr_scene_desc()
    C.pointer(bad)[0] = d
    bad.meshes = C.cast(meshes, C.POINTER(krr._abi.krr_mesh))
    lib = krr.load_library()
    wf = krr.WavefrontPathTracer(None)
    assert lib.krr_wf_set_scene(wf._h, C.byref(bad)) == krr._abi.KRR_ERR_INVALID_ARG
    assert b"triangle range" in lib.krr_last_error()
    wf.close()
    mk = krr.MegakernelPathTracer(None)
    assert lib.krr_mk_set_scene(mk._h, C.byref(bad)) == krr._abi.KRR_ERR_INVALID_ARG
    mk.close()
    h = C.c_void_p()
    assert lib.krr_accel_build(C.byref(bad), None, C.byref(h)) == krr._abi.KRR_ERR_INVALID_ARG


def test_noalloc_trace_argument_errors(krr):
    """The no-allocation trace checks its arguments before touching a device:
    a null accel or a negative count is KRR_ERR_INVALID_ARG with a message;
    its scratch array is 16 MiB."""
    lib = krr.load_library()
    assert lib.krr_accel_noalloc_scratch_bytes() == 16 << 20
    rc = lib.krr_accel_trace_closest_noalloc(None, None, 0, None, None)
    assert rc == -1 and b"noalloc" in lib.krr_last_error()
    rc = lib.krr_accel_trace_closest_noalloc(None, None, -5, None, None)
    assert rc == -1
