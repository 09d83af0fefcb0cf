#!/usr/bin/env python3
"""Benchmark of the wavefront path tracer on one MI355X.

Workload (default): 1024 icospheres (subdivision level 3, 1,310,720 triangles)
in the unit cube over a ground quad with 16 quad lights (1,310,754 triangles
in all), 1920x1080, 1 sample per pixel per frame, max depth 16, Russian
roulette from depth 3.  A step is one frame (beginFrame + render).

value = (closest-hit rays + shadow rays) / wall time of the timed frames, in
Mrays/s.  Rays are counted on the device (queue sizes summed over depth).
The roofline entry covers the kernel with the most HIP-event time per frame
(see kernel_rooflines for the byte model of each); its measured HBM traffic
comes from profiles/hbm_traffic_<config>.json when that file exists (rocprofv3
FETCH_SIZE + WRITE_SIZE per dispatch, scaled by a copy-kernel calibration).
cpu_baseline = the CPU checker (OpenMP) on full frames of the same scene.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s + Mpaths/s on 1 MI355X; equal-time relMSE: NRC vs plain PT"
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    "icosphere_field_1080p": dict(scene={"model": [{"type": "icosphere_field", "count": 1024, "level": 3, "seed": 7}],
                        "resolution": [1920, 1080],
                        "camera": {"eye": [0.5, 0.75, 2.3], "target": [0.5, 0.3, 0.45]}},
                 params={"spp": 1, "max_depth": 16, "rr": 0.8, "rr_min_depth": 3},
                 desc="icosphere_field_1.31Mtri_1920x1080_1spp_depth16_rr3",
                 # FP64 reference of the same view at 1/8 resolution, 4096 spp
                 reference_lowres=("tests/golden/icosphere_field_240x135_fp64_4096spp.npz", 240, 135),
                 # and at the measured resolution, 2048 spp
                 reference_fp64="tests/golden/icosphere_field_1920x1080_fp64_2048spp.npz"),
    "cornell_box_256": dict(scene={"model": [{"type": "cornell_box"}], "resolution": [256, 256],
                        "camera": {"eye": [0, 1, 2.8], "target": [0, 1, 0]}},
                 params={"spp": 16, "max_depth": 8, "rr": 0.8},
                 desc="cornell_box_36tri_256x256_16spp_depth8",
                 reference="tests/golden/cornell_256_fp64_65536spp.npz"),
    "cornell_box_1024": dict(scene={"model": [{"type": "cornell_box"}], "resolution": [1024, 1024],
                                    "camera": {"eye": [0, 1, 2.8], "target": [0, 1, 0]}},
                             params={"spp": 1, "max_depth": 8, "rr": 0.8},
                             desc="cornell_box_36tri_1024x1024_1spp_depth8",
                             reference_fp64="tests/golden/cornell_1024_fp64_1024spp.npz"),
    "volume_720p": dict(scene={"model": [{"type": "icosphere_field", "count": 1024, "level": 3, "seed": 7}],
                        "media": [{"type": "value_noise_grid", "res": 256, "octaves": 5, "seed": 11,
                                   "sigma_t_max": 20.0, "albedo": 0.8, "g": 0.3}],
                        "resolution": [1280, 720],
                        "camera": {"eye": [0.5, 0.75, 2.3], "target": [0.5, 0.3, 0.45], "medium": 0}},
                 params={"spp": 4, "max_depth": 32, "rr": 0.8, "rr_min_depth": 3},
                 desc="icosphere_field_1.31Mtri_volume256_1280x720_4spp_depth32"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gpu_bench(cfg, steps, warmup, traffic):
    import numpy as np
    import kiraray_amd as krr
    from kiraray_amd import hip

    if hip.device_count() < 1:
        raise SystemExit("bench.py: no HIP device visible")
    hip.check(hip.runtime().hipSetDevice(0))
    w, h = cfg["scene"]["resolution"]
    t0 = time.time()
    scene = krr.Scene(cfg["scene"])
    log(f"scene: {scene.n_triangles} triangles ({time.time() - t0:.1f}s)")
    params = dict(cfg["params"], time_kernels=True)
    wf = krr.WavefrontPathTracer(params)
    prm = wf.to_json()
    t0 = time.time()
    wf.set_scene(scene)
    log(f"set_scene (BVH build + upload): {time.time() - t0:.1f}s")
    # BVH build time of both builders (the same tree; the device one is the default)
    bvh = {"device_ms": round(wf.stats()["bvh_build_ms"], 2)}
    wh = krr.WavefrontPathTracer(dict(cfg["params"], bvh_builder="host"))
    wh.set_scene(scene)
    bvh["host_ms"] = round(wh.stats()["bvh_build_ms"], 2)
    wh.close()
    wd = krr.WavefrontPathTracer(dict(cfg["params"], bvh_builder="device"))
    wd.set_scene(scene)  # again, with the device code loaded and warm
    bvh["device_warm_ms"] = round(wd.stats()["bvh_build_ms"], 2)
    wd.close()
    log(f"bvh build: {bvh}")
    wf.resize(w, h)
    film = hip.DeviceBuffer(w * h * 16)
    stream = hip.Stream()
    cam = scene.camera(0)
    # launch shape (block size, LDS stack depth, occupancy) from the autotuner
    wf.begin_frame(cam, 1)
    t0 = time.time()
    tune = wf.autotune(stream=stream.handle)
    log(f"autotune {time.time() - t0:.2f}s: best {tune['best']}")
    prm = wf.to_json()

    def frame(idx):
        wf.begin_frame(cam, idx)
        wf.render(film.ptr, stream.handle)
        wf.end_frame()
        return wf.stats()

    for i in range(warmup):
        frame(1000 + i)
    stream.synchronize()
    hip.synchronize()
    ev0, ev1 = hip.Event(), hip.Event()
    tot = dict(closest=0, shadow=0, paths=0, trace_ms=0.0, shade_ms=0.0, shadow_ms=0.0, mega_ms=0.0, launches=0,
               mega_closest=0, mega_shadow=0)
    wall0 = time.perf_counter()
    ev0.record(stream)
    for i in range(steps):
        st = frame(1 + i)
        tot["closest"] += st["closest_rays"]
        tot["shadow"] += st["shadow_rays"]
        tot["paths"] += st["paths"]
        tot["trace_ms"] += st["closest_trace_ms"]
        tot["shade_ms"] += st["shade_ms"]
        tot["shadow_ms"] += st["shadow_trace_ms"]
        tot["mega_ms"] += st["megakernel_ms"]
        tot["launches"] += st["closest_launches"]
        tot["mega_closest"] += st["megakernel_closest_rays"]
        tot["mega_shadow"] += st["megakernel_shadow_rays"]
    ev1.record(stream)
    ev1.synchronize()
    stream.synchronize()
    wall = time.perf_counter() - wall0
    gpu_ms = ev0.elapsed_ms(ev1)
    log(f"timed {steps} frames: wall {wall * 1e3:.1f} ms, events {gpu_ms:.1f} ms")
    # Node / triangle counts of exactly these rays: same frame indices again,
    # with the device counters on (outside the timed region).
    wfc = krr.WavefrontPathTracer(dict(cfg["params"], count_traversal=True,
                                       **{k: prm[k] for k in ("trace_block_size", "stack_lds_entries", "occupancy")}))
    wfc.set_scene(scene)
    wfc.resize(w, h)
    keys = ("closest_rays", "shadow_rays", "node_visits", "tri_tests", "shadow_node_visits", "shadow_tri_tests",
            "megakernel_closest_rays", "megakernel_shadow_rays", "megakernel_node_visits", "megakernel_tri_tests")
    cnt = dict.fromkeys(keys, 0)
    for i in range(steps):
        wfc.begin_frame(cam, 1 + i)
        wfc.render(film.ptr, stream.handle)
        st = wfc.stats()
        for key in keys:
            cnt[key] += st[key]
    assert cnt["closest_rays"] == tot["closest"] and cnt["shadow_rays"] == tot["shadow"], \
        "counting pass traced different rays"
    img = film.to_numpy(np.float32, (h, w, 4))
    assert np.isfinite(img).all()
    relmse = None
    if cfg.get("reference"):
        # relMSE of the last timed frame against the FP64 reference image
        # (render/passes/errormeasure/metrics.cu:80-90), on the device
        z = np.load(os.path.join(ROOT, cfg["reference"]))
        ref = np.concatenate([z["image"], np.ones(z["image"].shape[:2] + (1,), np.float32)], -1)
        refb = hip.DeviceBuffer.from_numpy(np.ascontiguousarray(ref, np.float32))
        wf.begin_frame(cam, steps)
        wf.render(film.ptr, stream.handle)
        stream.synchronize()
        relmse = {"value": round(krr.error_metric(3, film.ptr, refb.ptr, w * h), 6),
                  "reference": f"{int(z['spp'])}-spp FP64 CPU oracle image ({cfg['reference']})",
                  "image": f"frame {steps}, {cfg['params']['spp']} spp"}
    if cfg.get("reference_lowres"):
        # per-image relMSE against the FP64 CPU oracle, at the reference's resolution
        path, lw, lh = cfg["reference_lowres"]
        z = np.load(os.path.join(ROOT, path))
        ref = np.concatenate([z["image"], np.ones(z["image"].shape[:2] + (1,), np.float32)], -1)
        refb = hip.DeviceBuffer.from_numpy(np.ascontiguousarray(ref, np.float32))
        wl = krr.WavefrontPathTracer(dict(cfg["params"]))
        wl.set_scene(scene)
        wl.resize(lw, lh)
        lf, la = hip.DeviceBuffer(lw * lh * 16), hip.DeviceBuffer(lw * lh * 32)
        single = []
        for f in range(16):
            wl.begin_frame(cam, 1 + f)
            wl.render(lf.ptr, stream.handle)
            stream.synchronize()
            single.append(krr.error_metric(3, lf.ptr, refb.ptr, lw * lh, stream.handle))
            krr.accumulate(la.ptr, lf.ptr, lw * lh, f, 0, 0, 64, stream.handle)
        stream.synchronize()
        relmse = {"value": round(float(np.median(single)), 5),
                  "reference": f"{int(z['spp'])}-spp FP64 CPU oracle image ({path})",
                  "image": f"median over frames 1..16 of single {cfg['params']['spp']}-spp frames at {lw}x{lh} (the 1080p view at 1/8 resolution)",
                  "running_16_frames": round(krr.error_metric(3, lf.ptr, refb.ptr, lw * lh, stream.handle), 5)}
        wl.close()
    wfc.close()
    wf.close()

    rays = tot["closest"] + tot["shadow"]
    secs = wall
    kernels = kernel_rooflines(cfg, prm, steps, tot, cnt)
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    k = kernels[dom]
    tr = (traffic or {}).get(k["kernel"])
    # The HBM roofline is the ceiling the traversal's bytes are priced
    # against; what limits the kernel is read from its own counters (profiles/):
    # measured HBM bytes over the launch time, L2 hit rate, VALU lane use.
    roofline = {"bound": "hbm", "achieved": k["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(k["achieved_gbs"] / HBM_PEAK_GBS, 4),
                "traffic": tr.get("bytes_per_launch") if tr else None,
                "kernel": k["kernel"], "avg_launch_ms": k["avg_launch_ms"],
                "algorithmic_bytes_per_launch": k["bytes_per_launch"],
                "bytes_model": k["bytes_model"], "per_ray": k["per_ray"]}
    if tr:
        roofline["traffic_source"] = tr.get("source")
        if k["avg_launch_ms"] > 0:
            measured_gbs = tr["bytes_per_launch"] / (k["avg_launch_ms"] * 1e-3) / 1e9
            roofline["measured_gbs"] = round(measured_gbs, 1)
            roofline["measured_frac"] = round(measured_gbs / HBM_PEAK_GBS, 4)
        for key in ("l2_hit_rate", "valu_lane_utilisation", "limiter"):
            if key in tr:
                roofline[key] = tr[key]
    out = {
        "value": round(rays / secs / 1e6, 2),
        "ms_per_step": round(secs * 1e3 / steps, 3),
        "mpaths_per_s": round(tot["paths"] / secs / 1e6, 2),
        "closest_mrays_per_s": round(tot["closest"] / secs / 1e6, 2),
        "rays_per_frame": rays // steps,
        "kernel_ms_per_step": {"closest_trace": round(tot["trace_ms"] / steps, 3),
                               "shade": round(tot["shade_ms"] / steps, 3),
                               "shadow_trace": round(tot["shadow_ms"] / steps, 3),
                               "megakernel": round(tot["mega_ms"] / steps, 3),
                               "gpu_events_total": round(gpu_ms / steps, 3)},
        "kernels": kernels,
        "roofline": roofline,
        "relmse": relmse,
        "autotune": {"best": tune["best"], "candidates": len(tune["candidates"]), "trial_paths": tune["trial_paths"]},
        "bvh_build": bvh,
        "_scene": scene,
    }
    return out


# Algorithmic bytes: what each kernel must move at least once per item, with
# no re-reads - a BVH4 node is one 128-byte line as floats (the default
# node_format) or 64 bytes quantised, a leaf triangle 48 bytes, a queued ray 32 bytes
# {o, tmax, d, meta} and its hit record 16 bytes.
TRI_B, RAY_B, HIT_B = 48, 32, 16
NODE_BYTES = {"quantized": 64, "float": 128}


def kernel_rooflines(cfg, prm, steps, tot, cnt):
    """Achieved algorithmic GB/s of each traversal kernel over the timed frames.

    prm: the pass's effective params (to_json).  HIP-event times of the
    launches (time_kernels) and device counts of the same frames
    (count_traversal, a second untimed pass)."""
    spp = prm["spp"]
    max_depth = prm["max_depth"]
    NODE_B = NODE_BYTES[prm.get("node_format", "float")]
    wdepths = prm["wavefront_depths"]
    mega_launches = steps * spp if wdepths <= max_depth else 0
    wf_closest = cnt["closest_rays"] - cnt["megakernel_closest_rays"]
    wf_shadow = cnt["shadow_rays"] - cnt["megakernel_shadow_rays"]
    out = {}

    def entry(name, kernel, ms, launches, byts, model, rays, nodes, tris):
        out[name] = {"kernel": kernel, "ms_per_step": round(ms / steps, 4),
                     "avg_launch_ms": round(ms / max(1, launches), 4), "launches_per_step": launches / steps,
                     "bytes_per_launch": int(byts / max(1, launches)), "bytes_model": model,
                     "achieved_gbs": round(byts / (ms * 1e-3) / 1e9, 2) if ms > 0 else 0.0,
                     "per_ray": {"node_visits": round(nodes / max(1, rays), 2),
                                 "tri_tests": round(tris / max(1, rays), 2)}}

    wf_launches = tot["launches"] - mega_launches
    entry("closest_trace", "k_trace_closest_queue", tot["trace_ms"], wf_launches,
          wf_closest * (RAY_B * 2 + HIT_B) + cnt["node_visits"] * NODE_B + cnt["tri_tests"] * TRI_B,
          f"64 B ray + 16 B hit per ray + {NODE_B} B per node visit + 48 B per triangle test",
          wf_closest, cnt["node_visits"], cnt["tri_tests"])
    shadow_launches = 0
    if prm["nee"]:  # no shadow rays are traced at the last depth
        shadow_launches = wf_launches - (steps * spp if max_depth < wdepths else 0)
    media = bool(cfg["scene"].get("media")) and prm["enable_medium"]
    entry("shadow_trace", "k_shadow_tr" if media else "k_trace_shadow", tot["shadow_ms"], shadow_launches,
          wf_shadow * (RAY_B * 2 + 16) + cnt["shadow_node_visits"] * NODE_B + cnt["shadow_tri_tests"] * TRI_B,
          f"64 B ray + 16 B contribution per ray + {NODE_B} B per node visit + 48 B per triangle test",
          wf_shadow, cnt["shadow_node_visits"], cnt["shadow_tri_tests"])
    mrays = cnt["megakernel_closest_rays"] + cnt["megakernel_shadow_rays"]
    refill = prm.get("refill_lanes", 0) > 0 and not media  # the persistent launch with dynamic ray refill
    entry("megakernel", "k_megakernel_refill" if refill else "k_megakernel", tot["mega_ms"], mega_launches,
          cnt["megakernel_node_visits"] * NODE_B + cnt["megakernel_tri_tests"] * TRI_B,
          f"traversal only (lower bound): {NODE_B} B per node visit + 48 B per triangle test, "
          "closest and shadow rays",
          mrays, cnt["megakernel_node_visits"], cnt["megakernel_tri_tests"])
    return out


NRC_PARAMS = {"train_mod": 16, "batch": 65536, "steps_per_frame": 4, "lr": 1e-3}


def load_reference_image(path, w, h):
    """FP64 checker image (film layout, row H-1-y first) as float32 RGBA, or
    None when the file is absent."""
    import numpy as np
    full = os.path.join(ROOT, path)
    if not os.path.exists(full):
        return None, None
    with np.load(full, allow_pickle=False) as z:
        img = z["image"].astype(np.float32)
        spp = int(z["spp"])
    assert img.shape == (h, w, 3), (img.shape, w, h)
    rgba = np.ones((h, w, 4), np.float32)
    rgba[..., :3] = img
    return rgba, spp


def nrc_equal_time(cfg, warm_frames=96, frames=24, ref_frames=96, ref_spp=4, nrc_params=None, tail_frames=96):
    """Radiance cache against the plain tracer on the same workload.

    Error reference: the FP64 checker image of the same view
    (cfg["reference_fp64"]); when that file is absent, the plain tracer with
    ref_frames x ref_spp samples per pixel (frame indices 100001..) stands
    in, and the output says so.  Plain: frames 1..frames.  Cache: warm_frames
    of online training, then frames more (training continues every frame).
    Each frame is timed on its own (host wall clock around begin_frame +
    render + sync, training included for the cache); the warm-up frames are
    timed too, so the wall-clock figures are given with and without them.
    Reports the relMSE of single frames, of the running average after equal
    wall time, the time each needs to reach plain's final relMSE, and the
    cache's bias: after tail_frames more frames, the relMSE floor of its
    running average and the relative error of its mean image brightness."""
    import numpy as np
    import kiraray_amd as krr
    from kiraray_amd import hip

    w, h = cfg["scene"]["resolution"]
    n = w * h
    scene = krr.Scene(cfg["scene"])
    cam = scene.camera(0)
    film = hip.DeviceBuffer(n * 16)
    acc = hip.DeviceBuffer(n * 32)
    ref = hip.DeviceBuffer(n * 16)
    stream = hip.Stream()

    def run(params, first, count, accumulate_into=None, warm=0, tail=0):
        wf = krr.WavefrontPathTracer(params)
        wf.set_scene(scene)
        wf.resize(w, h)
        stream.synchronize()
        t0 = time.perf_counter()
        for i in range(warm):
            wf.begin_frame(cam, first - warm + i)
            wf.render(film.ptr, stream.handle)
        stream.synchronize()
        t_warm = time.perf_counter() - t0
        times, single, running = [], [], []
        for i in range(count + tail):
            t0 = time.perf_counter()
            wf.begin_frame(cam, first + i)
            wf.render(film.ptr, stream.handle)
            stream.synchronize()
            times.append(time.perf_counter() - t0)
            if accumulate_into is None:
                if i < count:
                    single.append(krr.error_metric(3, film.ptr, ref.ptr, n, stream.handle))
                krr.accumulate(acc.ptr, film.ptr, n, i, 0, 0, 64, stream.handle)
                stream.synchronize()
                running.append(krr.error_metric(3, film.ptr, ref.ptr, n, stream.handle))
            else:
                krr.accumulate(accumulate_into.ptr, film.ptr, n, i, 0, 0, 64, stream.handle)
        stream.synchronize()
        st = wf.stats()
        mean_img = film.to_numpy(np.float32, (h, w, 4))[..., :3] if accumulate_into is None else None
        wf.close()
        return times, single, running, st, t_warm, mean_img

    base = dict(cfg["params"])
    ref_img, ref_spp_fp64 = load_reference_image(cfg["reference_fp64"], w, h) if cfg.get("reference_fp64") else (None, None)
    if ref_img is not None:
        log(f"nrc: FP64 checker reference {cfg['reference_fp64']} ({ref_spp_fp64} spp)")
        ref.copy_from_numpy(ref_img)
        ref_desc = f"FP64 CPU oracle, {ref_spp_fp64} spp ({cfg['reference_fp64']})"
    else:
        log(f"nrc: no FP64 reference image; plain tracer {ref_frames} x {ref_spp} spp instead")
        run(dict(base, spp=ref_spp), 100001, ref_frames, accumulate_into=acc)
        hip.check(hip.runtime().hipMemcpy(C_void(ref.ptr), C_void(film.ptr), n * 16, 3))
        ref_desc = f"plain tracer, {ref_frames * ref_spp} spp (no FP64 oracle image for this workload)"
    ref_mean = float(ref.to_numpy(np.float32, (h, w, 4))[..., :3].astype(np.float64).mean())
    log("nrc: plain tracer")
    tp, sp, rp, _, _, plain_img = run(base, 1, frames, tail=tail_frames)
    log("nrc: radiance cache")
    nrc_params = dict(NRC_PARAMS, **(nrc_params or {}))
    tn, sn, rn, stn, t_warm, nrc_img = run(dict(base, nrc=nrc_params, time_kernels=True), warm_frames + 1, frames, warm=warm_frames,
                                           tail=tail_frames)
    ms_p, ms_n = 1e3 * float(np.median(tp[:frames])), 1e3 * float(np.median(tn[:frames]))
    cum_p, cum_n = np.cumsum(tp[:frames]), np.cumsum(tn[:frames])
    rp_c, rn_c = rp[:frames], rn[:frames]
    # running average at equal wall time: the cache's curve at the time plain needs for all its frames
    t_plain = float(cum_p[-1])
    k = int(np.searchsorted(cum_n, t_plain, side="right"))
    rel_n_equal = rn_c[min(max(k, 1), len(rn_c)) - 1]
    # wall time each needs to bring its running average to plain's final relMSE
    target = float(rp_c[-1])
    hit = [i for i, x in enumerate(rn_c) if x <= target]
    t_nrc_reach = float(cum_n[hit[0]]) if hit else None
    # with the warm-up frames on the cache's clock
    t_nrc_reach_w = t_nrc_reach + t_warm if t_nrc_reach is not None else None

    def bias(img):
        return round((float(img.astype(np.float64).mean()) - ref_mean) / ref_mean, 5)

    return {"reference": ref_desc, "warm_frames": warm_frames, "frames": frames,
            "nrc_params": nrc_params,
            "plain": {"ms_per_frame": round(ms_p, 3), "relmse_single_frame": round(float(np.median(sp)), 5),
                      "relmse_running": [round(float(x), 5) for x in rp_c]},
            "nrc": {"ms_per_frame": round(ms_n, 3), "relmse_single_frame": round(float(np.median(sn)), 5),
                    "relmse_running": [round(float(x), 5) for x in rn_c], "loss": round(stn["nrc_loss"], 5),
                    "queries_per_frame": int(stn["nrc_queries"]), "train_vertices_per_frame": int(stn["nrc_train_vertices"]),
                    "kernel_ms_per_frame": round(stn["nrc_ms"], 3), "warm_up_s": round(t_warm, 4)},
            "equal_time": {"wall_s": round(t_plain, 4), "plain_relmse": round(float(rp_c[-1]), 5),
                           "nrc_frames": min(max(k, 1), len(rn_c)), "nrc_relmse": round(float(rel_n_equal), 5)},
            "time_to_plain_relmse": {"relmse": round(target, 5), "plain_s": round(t_plain, 4),
                                     "nrc_s": round(t_nrc_reach, 4) if t_nrc_reach else None,
                                     "nrc_frames": hit[0] + 1 if hit else None,
                                     "speedup": round(t_plain / t_nrc_reach, 3) if t_nrc_reach else None,
                                     "nrc_s_with_warm_up": round(t_nrc_reach_w, 4) if t_nrc_reach_w else None,
                                     "speedup_with_warm_up": round(t_plain / t_nrc_reach_w, 3) if t_nrc_reach_w else None},
            "converged": {"frames": frames + tail_frames,
                          "plain_relmse": round(float(rp[-1]), 5), "nrc_relmse": round(float(rn[-1]), 5),
                          "plain_mean_rel_error": bias(plain_img), "nrc_mean_rel_error": bias(nrc_img)},
            "speed_ratio_per_frame": round(ms_p / ms_n, 3),
            "single_frame_relmse_ratio": round(float(np.median(sp)) / float(np.median(sn)), 3)}


MULTIVIEW = dict(scene={"model": [{"type": "icosphere_field", "count": 1024, "level": 3, "seed": 7}],
                        "views": {"count": 256, "seed": 5, "radius": 2.0, "center": [0.5, 0.5, 0.5],
                                  "sizes": [128, 256, 512, 1024]}},
                 params={"spp": 2, "max_depth": 12, "rr": 0.8, "rr_min_depth": 3},
                 desc="icosphere_field_1.31Mtri_256views_128..1024sq_2spp_depth12")


def multiview_bench(frames=4, warmup=1, ref_frames=16, nrc_warm=8, nrc_params=None):
    """256 cameras on a radius-2 sphere around the icosphere field, square
    views of 128..1024 pixels, all rendered per frame by one persistent
    launch per sample (krr_wf_render_views), replayed from a recorded graph,
    launch shape from the autotuner.  Aggregate Mrays/s with the radiance
    cache off and on, and per-view relMSE of the last frame against a plain
    reference of ref_frames x 2 spp."""
    import numpy as np
    import kiraray_amd as krr
    from kiraray_amd import hip

    cfg = MULTIVIEW
    t0 = time.time()
    scene = krr.Scene(cfg["scene"])
    nv = scene.n_views
    sizes = [scene.view_size(v) for v in range(nv)]
    offs = np.concatenate([[0], np.cumsum([w * h for w, h in sizes])]).astype(np.int64)
    total = int(offs[-1])
    log(f"multiview: {nv} views, {total} pixels, scene {time.time() - t0:.1f}s")
    films = hip.DeviceBuffer(total * 16)  # all views' films back to back
    acc = hip.DeviceBuffer(total * 32)
    ref = hip.DeviceBuffer(total * 16)
    stream = hip.Stream()
    views = [(scene.camera(v), w, h, films.ptr + int(offs[v]) * 16) for v, (w, h) in enumerate(sizes)]

    def make(params):
        wf = krr.WavefrontPathTracer(params)
        t1 = time.time()
        wf.set_scene(scene)
        wf.set_views(views)
        return wf, time.time() - t1

    def timed(wf, first, count):
        rays = paths = 0
        stream.synchronize()
        w0 = time.perf_counter()
        per = []
        for i in range(count):
            t1 = time.perf_counter()
            wf.render_views(first + i, stream.handle)
            st = wf.stats()  # synchronises the stream
            per.append(time.perf_counter() - t1)
            rays += st["closest_rays"] + st["shadow_rays"]
            paths += st["paths"]
        return time.perf_counter() - w0, rays, paths, per, st

    def per_view_relmse():
        out = []
        for v, (w, h) in enumerate(sizes):
            o = int(offs[v]) * 16
            out.append(krr.error_metric(3, films.ptr + o, ref.ptr + o, w * h, stream.handle))
        a = np.array(out)
        return {"mean": round(float(a.mean()), 5), "median": round(float(np.median(a)), 5),
                "max": round(float(a.max()), 5), "min": round(float(a.min()), 5)}

    base = dict(cfg["params"])
    wf, t_set = make(base)
    t1 = time.time()
    tune = wf.autotune(stream=stream.handle)
    t_tune = time.time() - t1
    log(f"multiview: set_scene+views {t_set:.1f}s, autotune {t_tune:.2f}s best {tune['best']}")
    # reference: plain frames 100001.., accumulated
    for i in range(ref_frames):
        wf.render_views(100001 + i, stream.handle)
        krr.accumulate(acc.ptr, films.ptr, total, i, 0, 0, 64, stream.handle)
    stream.synchronize()
    hip.check(hip.runtime().hipMemcpy(C_void(ref.ptr), C_void(films.ptr), total * 16, 3))
    for i in range(warmup):
        wf.render_views(1000 + i, stream.handle)
    secs, rays, paths, per, st = timed(wf, 1, frames)
    plain_rel = per_view_relmse()
    eager, _ = make(dict(base, graph=False, **{k: tune["best"][k] for k in ("trace_block_size", "stack_lds_entries", "occupancy")}))
    eager.render_views(1000, stream.handle)
    e_secs, e_rays, _, _, _ = timed(eager, 1, 2)
    eager.close()
    plain = {"value": round(rays / secs / 1e6, 2), "ms_per_frame": round(secs * 1e3 / frames, 2),
             "mpaths_per_s": round(paths / secs / 1e6, 2), "rays_per_frame": rays // frames,
             "eager_ms_per_frame": round(e_secs * 1e3 / 2, 2), "relmse_per_view": plain_rel}
    prm = wf.to_json()
    wf.close()
    log(f"multiview plain: {plain}")
    nrc_params = dict(NRC_PARAMS, **(nrc_params or {}))
    wn, _ = make(dict(base, nrc=nrc_params, **{k: tune["best"][k] for k in ("trace_block_size", "stack_lds_entries", "occupancy")}))
    for i in range(nrc_warm):
        wn.render_views(2000 + i, stream.handle)
    n_secs, n_rays, n_paths, _, nst = timed(wn, 1 + nrc_warm, frames)
    nrc = {"value": round(n_rays / n_secs / 1e6, 2), "ms_per_frame": round(n_secs * 1e3 / frames, 2),
           "mpaths_per_s": round(n_paths / n_secs / 1e6, 2), "rays_per_frame": n_rays // frames,
           "relmse_per_view": per_view_relmse(), "queries_per_frame": int(nst["nrc_queries"]),
           "loss": round(nst["nrc_loss"], 5), "warm_frames": nrc_warm, "nrc_params": nrc_params}
    wn.close()
    log(f"multiview nrc: {nrc}")
    return {"workload": cfg["desc"], "views": nv, "pixels": total,
            "sizes": {str(s_): sum(1 for w, h in sizes if w == s_) for s_ in sorted(set(w for w, h in sizes))},
            "params": base, "graph": True, "autotune": {"seconds": round(t_tune, 2), "trial_paths": tune["trial_paths"],
                                                        "best": tune["best"], "candidates": len(tune["candidates"])},
            "launch": {k: prm[k] for k in ("trace_block_size", "stack_lds_entries", "occupancy")},
            "reference": f"plain tracer, {ref_frames * base['spp']} spp per pixel",
            "plain": plain, "nrc": nrc}


def node_format_ab(cfg, frames=10, warmup=3):
    """The benchmark workload with each BVH node format - 128-byte float
    nodes (the default) and 64-byte quantised nodes - in one process: ms per
    frame and Mrays/s (HIP events on the render stream), node visits and
    node bytes per ray (count_traversal frames).  Each side gets its own
    autotuned launch shape.  The images are the same (the closest hit does
    not depend on the tree's boxes)."""
    import kiraray_amd as krr
    from kiraray_amd import hip

    w, h = cfg["scene"]["resolution"]
    scene = krr.Scene(cfg["scene"])
    cam = scene.camera(0)
    film = hip.DeviceBuffer(w * h * 16)
    stream = hip.Stream()
    out = {"workload": cfg["desc"], "frames": frames}
    for fmt in ("quantized", "float"):
        wf = krr.WavefrontPathTracer(dict(cfg["params"], node_format=fmt))
        wf.set_scene(scene)
        wf.resize(w, h)
        wf.begin_frame(cam, 1)
        wf.autotune(stream=stream.handle)
        for i in range(warmup):
            wf.begin_frame(cam, 500 + i)
            wf.render(film.ptr, stream.handle)
        ev0, ev1 = hip.Event(), hip.Event()
        rays = 0
        ev0.record(stream)
        for i in range(frames):
            wf.begin_frame(cam, 1 + i)
            wf.render(film.ptr, stream.handle)
            st = wf.stats()
            rays += st["closest_rays"] + st["shadow_rays"]
        ev1.record(stream)
        ev1.synchronize()
        ms = ev0.elapsed_ms(ev1) / frames
        wf.close()
        wc = krr.WavefrontPathTracer(dict(cfg["params"], node_format=fmt, count_traversal=True))
        wc.set_scene(scene)
        wc.resize(w, h)
        wc.begin_frame(cam, 1)
        wc.render(film.ptr, stream.handle)
        st = wc.stats()
        wc.close()
        nr = st["megakernel_closest_rays"] + st["megakernel_shadow_rays"]
        visits = st["megakernel_node_visits"] / max(1, nr)
        out[fmt] = {"ms_per_frame": round(ms, 3), "mrays_per_s": round(rays / frames / ms / 1e3, 1),
                    "node_visits_per_ray": round(visits, 2), "node_bytes_per_ray": round(visits * NODE_BYTES[fmt], 1)}
        log(f"node format {fmt}: {out[fmt]}")
    return out


def refill_ab(cfg, values=(0, 2, 4, 6, 8, 12, 16, 32), frames=10, warmup=3):
    """The benchmark workload with the persistent launch's refill threshold
    (pass param refill_lanes: a wave's traversal round ends once fewer lanes
    than this are still traversing; 0 runs every ray to its end)."""
    return param_sweep(cfg, "refill_lanes", values, frames, warmup)


def param_sweep(cfg, key, values, frames=10, warmup=3):
    """The benchmark workload with each value of one pass param in one
    process: ms per frame and Mrays/s (HIP events on the render stream, no
    host synchronisation inside the timed frames), each value with its own
    autotuned launch shape.  For launch params the images do not depend on
    the value."""
    import kiraray_amd as krr
    from kiraray_amd import hip

    w, h = cfg["scene"]["resolution"]
    scene = krr.Scene(cfg["scene"])
    cam = scene.camera(0)
    film = hip.DeviceBuffer(w * h * 16)
    stream = hip.Stream()
    out = {"workload": cfg["desc"], "frames": frames, "param": key}
    for v in values:
        wf = krr.WavefrontPathTracer(dict(cfg["params"], **{key: v}))
        wf.set_scene(scene)
        wf.resize(w, h)
        wf.begin_frame(cam, 1)
        tune = wf.autotune(stream=stream.handle)
        for i in range(warmup):
            wf.begin_frame(cam, 500 + i)
            wf.render(film.ptr, stream.handle)
        ev0, ev1 = hip.Event(), hip.Event()
        ev0.record(stream)
        for i in range(frames):
            wf.begin_frame(cam, 1 + i)
            wf.render(film.ptr, stream.handle)
        ev1.record(stream)
        ev1.synchronize()
        ms = ev0.elapsed_ms(ev1) / frames
        st = wf.stats()  # the last timed frame
        rays = st["closest_rays"] + st["shadow_rays"]
        wf.close()
        out[str(v)] = {"ms_per_frame": round(ms, 3), "mrays_per_s": round(rays / ms / 1e3, 1),
                       "launch": {k: tune["best"][k] for k in ("trace_block_size", "stack_lds_entries", "occupancy")}}
        log(f"{key} {v}: {out[str(v)]}")
    return out


def megakernel_ab(cfg, frames=5, warmup=2):
    """The megakernel path tracer (one lane per pixel, MegakernelPathTracer)
    against the wavefront pass in its default form - every depth of every
    path in one persistent launch whose lanes take a new camera path as soon
    as theirs ends ("persistent_lane") - on the same scene, view, spp and max depth,
    both with the reference's Russian roulette (rr 0.8 from the first vertex,
    as MegakernelPathTracer has it): ms per frame and Mrays/s of each.  Each
    side gets its best launch shape (the wavefront pass from its autotuner,
    the megakernel from block sizes 64 / 128 / 256)."""
    import kiraray_amd as krr
    from kiraray_amd import hip

    w, h = cfg["scene"]["resolution"]
    scene = krr.Scene(cfg["scene"])
    cam = scene.camera(0)
    base = {"spp": cfg["params"]["spp"], "max_depth": cfg["params"]["max_depth"], "rr": 0.8}
    film = hip.DeviceBuffer(w * h * 16)
    stream = hip.Stream()

    def run(p, first, count):
        st_sum = dict(closest=0, shadow=0, paths=0)
        ev0, ev1 = hip.Event(), hip.Event()
        ev0.record(stream)
        for i in range(count):
            p.begin_frame(cam, first + i)
            p.render(film.ptr, stream.handle)
            p.end_frame()
            st = p.stats()
            st_sum["closest"] += st["closest_rays"]
            st_sum["shadow"] += st["shadow_rays"]
            st_sum["paths"] += st["paths"]
        ev1.record(stream)
        ev1.synchronize()
        ms = ev0.elapsed_ms(ev1) / count
        rays = (st_sum["closest"] + st_sum["shadow"]) / count
        return {"ms_per_frame": round(ms, 3), "mrays_per_s": round(rays / ms / 1e3, 1),
                "mpaths_per_s": round(st_sum["paths"] / count / ms / 1e3, 1), "rays_per_frame": int(rays)}

    out = {"workload": cfg["desc"].replace("_rr3", "") + "_rr0.8_from_depth0", "frames": frames}
    wf = krr.WavefrontPathTracer(dict(base, rr_min_depth=0))
    wf.set_scene(scene)
    wf.resize(w, h)
    wf.begin_frame(cam, 1)
    wf.autotune(stream=stream.handle)
    run(wf, 500, warmup)
    out["persistent_lane"] = run(wf, 1, frames)
    wf.close()
    best = None
    for tb in (64, 128, 256):
        mk = krr.MegakernelPathTracer(dict(base, block_size=tb))
        mk.set_scene(scene)
        mk.resize(w, h)
        run(mk, 500, 1)
        r = run(mk, 1, max(2, frames // 2))
        log(f"megakernel block {tb}: {r}")
        if best is None or r["ms_per_frame"] < best[1]["ms_per_frame"]:
            best = (tb, r)
        mk.close()
    mk = krr.MegakernelPathTracer(dict(base, block_size=best[0]))
    mk.set_scene(scene)
    mk.resize(w, h)
    run(mk, 500, warmup)
    out["megakernel"] = dict(run(mk, 1, frames), block_size=best[0])
    mk.close()
    out["persistent_lane_speedup"] = round(out["megakernel"]["ms_per_frame"] / out["persistent_lane"]["ms_per_frame"], 2)
    log(f"megakernel A/B: {out}")
    return out


def C_void(p):
    import ctypes
    return ctypes.c_void_p(p)


def cpu_baseline(cfg, scene, budget_s=12.0, max_frames=20):
    import oracle.binding as ob
    w, h = cfg["scene"]["resolution"]
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    rays = paths = 0
    secs = 0.0
    frames = 0
    cam = scene.camera(0)
    prm = dict(cfg["params"])
    kw = {k: int(v) if isinstance(v, bool) else v for k, v in prm.items()
          if k in ("spp", "max_depth", "rr", "rr_min_depth")}
    while frames < max_frames and secs < budget_s:
        _, st = ob.render(scene.desc, cam, w, h, 1 + frames, n_threads=cores, **kw)
        rays += st["closest_rays"] + st["shadow_rays"]
        paths += st["paths"]
        secs += st["seconds"]
        frames += 1
        cores_used = st["threads"]
    return {"value": round(rays / secs / 1e6, 3), "unit": "Mrays/s", "cores": cores_used, "kind": "port",
            "sample": f"{frames} full frame(s) {w}x{h} of the same scene and params (frames 1..{frames}), "
                      f"{paths} paths, {secs:.1f}s of render time, BVH build excluded",
            "mpaths_per_s": round(paths / secs / 1e6, 3)}


def copy_calibration(nbytes=1 << 30, reps=3):
    """A plain device-to-device copy of known size (torch's elementwise copy
    kernel), run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE: the counters
    should read nbytes each; the measured ratio scales the tracer kernels'
    counter readings into HBM bytes."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    assert bool((b == 1.0).all())
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="icosphere_field_1080p", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-nrc", action="store_true", help="skip the radiance-cache equal-time comparison")
    ap.add_argument("--no-multiview", action="store_true", help="skip the 256-view batched frame")
    ap.add_argument("--no-megakernel", action="store_true", help="skip the megakernel path tracer A/B")
    ap.add_argument("--section", default="all", choices=["all", "nrc", "multiview", "megakernel", "nodes", "refill", "sweep"],
                    help="nrc / multiview: run only that comparison and print its JSON (profiling runs)")
    ap.add_argument("--nrc-warm", type=int, default=96, help="training frames before the NRC comparison")
    ap.add_argument("--nrc-frames", type=int, default=24, help="compared frames of the NRC comparison")
    ap.add_argument("--nrc-tail", type=int, default=96,
                    help="further frames of the NRC comparison for the converged relMSE and bias (0 for profiling runs)")
    ap.add_argument("--wavefront-depths", type=int, default=None,
                    help="run the first N depths of each path as per-depth queue kernels (the pass param wavefront_depths)")
    ap.add_argument("--sweep-key", default="lds_nodes", help="--section sweep: the pass param to vary")
    ap.add_argument("--sweep-values", default="0,21,85,341", help="--section sweep: comma-separated integer values")
    ap.add_argument("--refill-lanes", type=int, default=None,
                    help="the pass param refill_lanes of the benchmark workload (default: the pass default)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: runs on exactly one GPU (the path shards by image rows or views with one film exchange per "
              "frame; that sharding is not launched here); use --gpus 1", file=sys.stderr)
        raise SystemExit(2)
    cfg = CONFIGS[args.config]
    if args.wavefront_depths is not None:
        cfg = dict(cfg, params=dict(cfg["params"], wavefront_depths=args.wavefront_depths))
    if args.refill_lanes is not None:
        cfg = dict(cfg, params=dict(cfg["params"], refill_lanes=args.refill_lanes))
    if args.section != "all":
        from kiraray_amd import hip
        hip.check(hip.runtime().hipSetDevice(0))
        if args.section == "nrc":
            out = {"icosphere_field_1080p": nrc_equal_time(CONFIGS["icosphere_field_1080p"], warm_frames=args.nrc_warm,
                                                           frames=args.nrc_frames, tail_frames=args.nrc_tail)}
        elif args.section == "megakernel":
            out = megakernel_ab(CONFIGS["icosphere_field_1080p"])
        elif args.section == "nodes":
            out = node_format_ab(cfg)
        elif args.section == "refill":
            out = refill_ab(cfg)
        elif args.section == "sweep":
            out = param_sweep(cfg, args.sweep_key, [int(v) for v in args.sweep_values.split(",")])
        else:
            out = multiview_bench()
        print(json.dumps(out), flush=True)
        return
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"hbm_traffic_{args.config}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)
    r = gpu_bench(cfg, args.steps, args.warmup, traffic)
    scene = r.pop("_scene")
    line = {"metric": METRIC, "value": r["value"], "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic (procedural seeded scene)",
            "config": {"workload": cfg["desc"], "triangles": scene.n_triangles,
                       "resolution": cfg["scene"]["resolution"], **cfg["params"]},
            "mpaths_per_s": r["mpaths_per_s"], "closest_mrays_per_s": r["closest_mrays_per_s"],
            "rays_per_frame": r["rays_per_frame"], "kernel_ms_per_step": r["kernel_ms_per_step"],
            "kernels": r["kernels"],
            "relmse": r["relmse"], "autotune": r["autotune"], "bvh_build": r["bvh_build"],
            "roofline": r["roofline"]}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(cfg, scene, budget_s=args.cpu_budget)
    else:
        line["cpu_baseline"] = None
    if args.config == "icosphere_field_1080p" and not args.no_nrc:
        # the open workload, and a closed scene where paths are long
        line["nrc_vs_plain"] = {"icosphere_field_1080p": nrc_equal_time(cfg, warm_frames=args.nrc_warm, frames=args.nrc_frames,
                                                                        tail_frames=args.nrc_tail),
                                "cornell_box_1024": nrc_equal_time(CONFIGS["cornell_box_1024"],
                                                                   nrc_params={"steps_per_frame": 2, "batch": 32768},
                                                                   tail_frames=args.nrc_tail)}
    if args.config == "icosphere_field_1080p" and not args.no_multiview:
        line["multiview"] = multiview_bench()
    if args.config == "icosphere_field_1080p" and not args.no_megakernel:
        line["megakernel_ab"] = megakernel_ab(cfg)
        line["node_format_ab"] = node_format_ab(cfg)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
