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

usage: python tests/golden/make_icosphere_reference.py [--spp-per-chunk N] [--chunks K]
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
W, H = 240, 135


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--spp-per-chunk", type=int, default=256)
    ap.add_argument("--chunks", type=int, default=16)
    ap.add_argument("--out", default=ICOSPHERE_REFERENCE)
    args = ap.parse_args()
    import kiraray_amd as krr
    import oracle.binding as ob
    from conftest import icosphere_scene
    s = krr.Scene(icosphere_scene())  # 1920x1080 config: the camera's aspect ratio is 16:9, as at 240x135
    acc = np.zeros((H, W, 3), np.float64)
    t0 = time.time()
    for c in range(args.chunks):
        img, st = ob.render(s.desc, s.camera(0), W, H, 2001 + c, spp=args.spp_per_chunk, max_depth=16, rr=0.8, rr_min_depth=3,
                            precision=64)
        acc += img[..., :3].astype(np.float64)
        print(f"chunk {c + 1}/{args.chunks}: {st['seconds']:.1f}s on {st['threads']} threads "
              f"({time.time() - t0:.0f}s total)", flush=True)
    ref = (acc / args.chunks).astype(np.float32)
    np.savez_compressed(args.out, image=ref, spp=np.int64(args.spp_per_chunk * args.chunks),
                        frames=np.arange(2001, 2001 + args.chunks), max_depth=np.int64(16), rr_min_depth=np.int64(3))
    print(f"wrote {args.out}: mean {ref.mean():.6f}", flush=True)


if __name__ == "__main__":
    main()
