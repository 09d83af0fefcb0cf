"""Many views in one frame (krr_wf_set_views / krr_wf_render_views).

The multi-view frame has no reference counterpart; it batches the
reference's per-view frame (integrator.cpp:205-267).  What pins it:
* view 0 seeds exactly as the single-view pass, so a one-view frame must be
  bit-identical to krr_wf_render of the same camera and size;
* view v > 0 seeds with initstate interleave(x, y) | v << 32, which the CPU
  checker restates (view_id), so every view of an uneven batch must match
  the checker's image of that view and the ray counts must add up;
* the recorded graph, replayed frame after frame with eager work in
  between, must give the eager result of every frame;
* the autotuner may only change launch shapes, never the image.
"""
import numpy as np
import pytest

from conftest import CORNELL, icosphere_scene

pytestmark = pytest.mark.gpu


def _films(gpu, sizes):
    return [gpu.DeviceBuffer(w * h * 16) for (w, h) in sizes]


def _read(gpu, films, sizes):
    gpu.synchronize()
    return [f.to_numpy(np.float32, (h, w, 4)) for f, (w, h) in zip(films, sizes)]


def _assert_match(g, o, rtol=1e-4, atol=1e-6):
    d = np.abs(g[..., :3] - o[..., :3])
    bad = (d > rtol * np.abs(o[..., :3]) + atol).any(-1)
    print(f"pixels {g.shape[0] * g.shape[1]} exact {int((g == o).all(-1).sum())} outside 1e-4: {int(bad.sum())}", flush=True)
    assert int(bad.sum()) == 0
    assert np.all(g[..., 3] == 1.0)


@pytest.mark.parametrize("graph", [False, True])
def test_one_view_equals_single_view_render(krr, gpu, graph):
    """Cornell 96x80, 2 spp, depth 8, frame 7: render_views with one view ==
    begin_frame + render, bit for bit, with equal ray counts."""
    print(f"case: views one view == render graph={graph}", flush=True)
    s = krr.Scene(CORNELL)
    params = {"spp": 2, "max_depth": 8, "graph": graph}
    cam = s.camera(0)
    wf = krr.WavefrontPathTracer(params)
    wf.set_scene(s)
    wf.resize(96, 80)
    film = gpu.DeviceBuffer(96 * 80 * 16)
    wf.begin_frame(cam, 7)
    wf.render(film.ptr, 0)
    st1 = wf.stats()
    ref = film.to_numpy(np.float32, (80, 96, 4))
    vfilm = gpu.DeviceBuffer(96 * 80 * 16)
    wf.set_views([(cam, 96, 80, vfilm.ptr)])
    wf.render_views(7, 0)
    st2 = wf.stats()
    img = vfilm.to_numpy(np.float32, (80, 96, 4))
    wf.close()
    print(st1["closest_rays"], st2["closest_rays"], st1["shadow_rays"], st2["shadow_rays"])
    assert st1["closest_rays"] == st2["closest_rays"] and st1["shadow_rays"] == st2["shadow_rays"]
    assert st2["paths"] == 96 * 80 * 2
    assert np.array_equal(img, ref)


def test_uneven_views_match_oracle(krr, oracle, gpu):
    """Four views of a 20,480-triangle icosphere field at different square
    sizes, 2 spp, depth 12, RR from depth 3: each view equals the checker's
    image of that view (view_id seeding), the ray counts sum up."""
    print("case: views uneven 4 views vs oracle", flush=True)
    cfg = icosphere_scene(count=64, level=2, res=(64, 64))
    cfg["views"] = {"count": 4, "seed": 3, "radius": 2.0, "sizes": [24, 40, 56]}
    s = krr.Scene(cfg)
    n = s.n_views
    sizes = [s.view_size(v) for v in range(n)]
    print("sizes", sizes, flush=True)
    assert len(set(sizes)) > 1
    params = {"spp": 2, "max_depth": 12, "rr_min_depth": 3}
    wf = krr.WavefrontPathTracer(params)
    wf.set_scene(s)
    films = _films(gpu, sizes)
    wf.set_views([(s.camera(v), w, h, f.ptr) for v, ((w, h), f) in enumerate(zip(sizes, films))])
    wf.render_views(3, 0)
    st = wf.stats()
    imgs = _read(gpu, films, sizes)
    wf.close()
    cr = sr = 0
    for v, (w, h) in enumerate(sizes):
        o, os_ = oracle.render(s.desc, s.camera(v), w, h, 3, spp=2, max_depth=12, rr_min_depth=3, view_id=v)
        cr += os_["closest_rays"]
        sr += os_["shadow_rays"]
        _assert_match(imgs[v], o)
    print("gpu", st["closest_rays"], st["shadow_rays"], "cpu", cr, sr)
    assert st["closest_rays"] == cr and st["shadow_rays"] == sr
    assert st["paths"] == 2 * sum(w * h for w, h in sizes)


def test_view_ids_decorrelate(krr, gpu):
    """Two views with the same camera and size but different view ids must
    get different noise (different seeds), with the same mean."""
    print("case: views decorrelate", flush=True)
    s = krr.Scene(CORNELL)
    cam = s.camera(0)
    wf = krr.WavefrontPathTracer({"spp": 1, "max_depth": 8})
    wf.set_scene(s)
    films = _films(gpu, [(64, 64)] * 2)
    wf.set_views([(cam, 64, 64, films[0].ptr), (cam, 64, 64, films[1].ptr)])
    wf.render_views(1, 0)
    a, b = _read(gpu, films, [(64, 64)] * 2)
    wf.close()
    assert not np.array_equal(a, b)
    assert abs(a[..., :3].mean() - b[..., :3].mean()) < 0.1 * a[..., :3].mean()


def test_graph_replays_match_eager(krr, gpu):
    """The recorded frame, replayed for frames 1..4 with eager launches and
    synchronisations in between, equals an eager frame of the same index
    every time; different frame indices give different images."""
    print("case: views graph replays vs eager", flush=True)
    cfg = icosphere_scene(count=64, level=2, res=(64, 64))
    cfg["views"] = {"count": 5, "seed": 9, "radius": 2.0, "sizes": [16, 32, 48]}
    s = krr.Scene(cfg)
    sizes = [s.view_size(v) for v in range(s.n_views)]
    base = {"spp": 2, "max_depth": 12, "rr_min_depth": 3}
    wg = krr.WavefrontPathTracer({**base, "graph": True})
    we = krr.WavefrontPathTracer({**base, "graph": False})
    fg, fe = _films(gpu, sizes), _films(gpu, sizes)
    for wf, films in ((wg, fg), (we, fe)):
        wf.set_scene(s)
        wf.set_views([(s.camera(v), w, h, f.ptr) for v, ((w, h), f) in enumerate(zip(sizes, films))])
    prev = None
    for frame in (1, 2, 3, 4):
        wg.render_views(frame, 0)
        gpu.synchronize()
        we.render_views(frame, 0)
        g = _read(gpu, fg, sizes)
        e = _read(gpu, fe, sizes)
        sg, se = wg.stats(), we.stats()
        print(f"frame {frame} rays graph {sg['closest_rays']} eager {se['closest_rays']}", flush=True)
        assert sg["closest_rays"] == se["closest_rays"] and sg["shadow_rays"] == se["shadow_rays"]
        for a, b in zip(g, e):
            assert np.array_equal(a, b)
        if prev is not None:
            assert not all(np.array_equal(a, b) for a, b in zip(g, prev))
        prev = g
    wg.close()
    we.close()


def test_graph_with_radiance_cache(krr, gpu):
    """NRC in the recorded frame: frame 1 (initial weights) equals the eager
    frame bit for bit; every replay takes its Adam steps (device-side step
    count), answers queries and records training vertices."""
    print("case: views graph + nrc", flush=True)
    s = krr.Scene(CORNELL)
    cam = s.camera(0)
    sizes = [(64, 64), (48, 48)]
    nrc = {"train_mod": 4, "batch": 8192, "steps_per_frame": 2, "lr": 3e-3}
    base = {"spp": 1, "max_depth": 8, "nrc": nrc}
    wg = krr.WavefrontPathTracer({**base, "graph": True})
    we = krr.WavefrontPathTracer({**base, "graph": False})
    fg, fe = _films(gpu, sizes), _films(gpu, sizes)
    for wf, films in ((wg, fg), (we, fe)):
        wf.set_scene(s)
        wf.set_views([(cam, w, h, f.ptr) for (w, h), f in zip(sizes, films)])
    for frame in range(1, 7):
        wg.render_views(frame, 0)
        we.render_views(frame, 0)
        g = _read(gpu, fg, sizes)
        e = _read(gpu, fe, sizes)
        st = wg.stats()
        print(f"frame {frame} queries {st['nrc_queries']} train {st['nrc_train_vertices']} steps {st['nrc_steps']} "
              f"loss {st['nrc_loss']:.4f}", flush=True)
        if frame == 1:
            for a, b in zip(g, e):
                assert np.array_equal(a, b)
        for a in g:
            assert np.isfinite(a).all() and (a[..., :3] >= 0).all()
        assert st["nrc_steps"] == 2 * frame
        assert st["nrc_queries"] > 0 and st["nrc_train_vertices"] > 0
    wg.close()
    we.close()


@pytest.mark.parametrize("views", [False, True])
def test_autotune_keeps_the_image(krr, gpu, views):
    """The autotuner times 27 launch shapes, then the four other refill
    thresholds of {0, 4, 8, 12, 16} at the fastest shape (refill_lanes left to
    the scene), and keeps the fastest; the frame rendered afterwards equals
    the frame rendered before."""
    print(f"case: autotune views={views}", flush=True)
    cfg = icosphere_scene(count=64, level=2, res=(96, 64))
    s = krr.Scene(cfg)
    cam = s.camera(0)
    wf = krr.WavefrontPathTracer({"spp": 1, "max_depth": 12, "rr_min_depth": 3})
    wf.set_scene(s)
    film = gpu.DeviceBuffer(96 * 64 * 16)

    def frame():
        if views:
            wf.render_views(2, 0)
        else:
            wf.begin_frame(cam, 2)
            wf.render(film.ptr, 0)
        gpu.synchronize()
        return film.to_numpy(np.float32, (64, 96, 4))

    if views:
        wf.set_views([(cam, 96, 64, film.ptr)])
    else:
        wf.resize(96, 64)
    before = frame()
    rep = wf.autotune(trial_paths=96 * 64)
    print("best", rep["best"], flush=True)
    assert len(rep["candidates"]) == 31
    assert sorted({c["refill_lanes"] for c in rep["candidates"]}) == [0, 4, 8, 12, 16]
    assert all(c["ms"] > 0 for c in rep["candidates"])
    assert rep["best"]["ms"] == min(c["ms"] for c in rep["candidates"])
    p = wf.to_json()
    assert p["trace_block_size"] == rep["best"]["trace_block_size"]
    assert p["stack_lds_entries"] == rep["best"]["stack_lds_entries"]
    assert p["refill_lanes"] == p["refill_lanes_in_force"] == rep["best"]["refill_lanes"]
    after = frame()
    wf.close()
    assert np.array_equal(before, after)
