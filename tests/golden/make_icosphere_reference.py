"""Writes tests/golden/icosphere_field_240x135_fp64_4096spp.npz: the
1.31 M-triangle icosphere field (1024 level-3 icospheres, seed 7, the
benchmark's camera) at 240x135 - the 1920x1080 frame's view at 1/8
resolution - max depth 16, RR 0.8 from depth 3, rendered by the CPU checker
in FP64 (reference-image mode) with 4,096 samples per pixel, as 16
independent frames of 256 spp (frame indices 2001..2016) averaged in
float64.

It is the FP64 error reference of per-image relMSE reports on this scene
(the reference's metric, src/render/passes/errormeasure/metrics.cu:80-90).
A 1080p image at 65,536 spp would take hours of CPU time; this one takes
about a minute on 8 cores.  Data produced by this repository's own checker.

With --full: tests/golden/icosphere_field_1920x1080_fp64_2048spp.npz, the
same view at the measured 1920x1080, 2,048 spp (32 frames of 64 spp, frame
indices 2001..2032), stored as float16 (relative rounding 2^-11, far below
the relMSE differences it measures); about 35 minutes on 8 cores.  It is
the FP64 reference of the radiance cache's equal-time comparison.

usage: python tests/golden/make_icosphere_reference.py [--full] [--spp-per-chunk N] [--chunks K]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

ICOSPHERE_REFERENCE = os.path.join(HERE, "icosphere_field_240x135_fp64_4096spp.npz")
ICOSPHERE_REFERENCE_1080P = os.path.join(HERE, "icosphere_field_1920x1080_fp64_2048spp.npz")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--full", action="store_true", help="1920x1080, 32 x 64 spp, float16")
    ap.add_argument("--spp-per-chunk", type=int, default=None)
    ap.add_argument("--chunks", type=int, default=None)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    W, H = (1920, 1080) if args.full else (240, 135)
    spc = args.spp_per_chunk or (64 if args.full else 256)
    chunks = args.chunks or (32 if args.full else 16)
    out = args.out or (ICOSPHERE_REFERENCE_1080P if args.full else ICOSPHERE_REFERENCE)
    import kiraray_amd as krr
    import oracle.binding as ob
    from conftest import icosphere_scene
    s = krr.Scene(icosphere_scene())  # 1920x1080 config: the camera's aspect ratio is 16:9, as at 240x135
    acc = np.zeros((H, W, 3), np.float64)
    t0 = time.time()
    for c in range(chunks):
        img, st = ob.render(s.desc, s.camera(0), W, H, 2001 + c, spp=spc, max_depth=16, rr=0.8, rr_min_depth=3,
                            precision=64)
        acc += img[..., :3].astype(np.float64)
        print(f"chunk {c + 1}/{chunks}: {st['seconds']:.1f}s on {st['threads']} threads "
              f"({time.time() - t0:.0f}s total)", flush=True)
    ref = (acc / chunks).astype(np.float16 if args.full else np.float32)
    np.savez_compressed(out, image=ref, spp=np.int64(spc * chunks),
                        frames=np.arange(2001, 2001 + chunks), max_depth=np.int64(16), rr_min_depth=np.int64(3))
    print(f"wrote {out}: mean {ref.astype(np.float64).mean():.6f}", flush=True)


if __name__ == "__main__":
    main()
