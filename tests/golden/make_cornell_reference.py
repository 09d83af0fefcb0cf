"""Writes tests/golden/cornell_256_fp64_65536spp.npz: the Cornell box
(256x256, max depth 8, NEE, RR 0.8) rendered by the CPU checker in FP64
(reference-image mode) with 65,536 samples per pixel, as 16 independent
frames of 4096 spp (frame indices 1001..1016) averaged in float64.

The image is the error reference for relMSE reports (the reference's
metric, src/render/passes/errormeasure/metrics.cu:80-90).  It is data
produced by this repository's own checker, not by the reference.

With --res1024: tests/golden/cornell_1024_fp64_1024spp.npz, the same box at
1024x1024 (max depth 8, NEE, RR 0.8), 1,024 spp as 16 frames of 64 spp
(frame indices 1001..1016), stored as float16; the FP64 reference of the
radiance cache's equal-time comparison on the closed scene.

usage: python tests/golden/make_cornell_reference.py [--res1024] [--spp-per-chunk N] [--chunks K]
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

CORNELL_REFERENCE = os.path.join(HERE, "cornell_256_fp64_65536spp.npz")
CORNELL_REFERENCE_1024 = os.path.join(HERE, "cornell_1024_fp64_1024spp.npz")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--res1024", action="store_true", help="1024x1024, 16 x 64 spp, float16")
    ap.add_argument("--spp-per-chunk", type=int, default=None)
    ap.add_argument("--chunks", type=int, default=16)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    import kiraray_amd as krr
    import oracle.binding as ob
    from conftest import CORNELL
    res = 1024 if args.res1024 else 256
    spc = args.spp_per_chunk or (64 if args.res1024 else 4096)
    out = args.out or (CORNELL_REFERENCE_1024 if args.res1024 else CORNELL_REFERENCE)
    s = krr.Scene(dict(CORNELL, resolution=[res, res]))
    acc = np.zeros((res, res, 3), np.float64)
    t0 = time.time()
    for c in range(args.chunks):
        img, st = ob.render(s.desc, s.camera(0), res, res, 1001 + c, spp=spc, max_depth=8, nee=1, rr=0.8, precision=64)
        acc += img[..., :3].astype(np.float64)
        print(f"chunk {c + 1}/{args.chunks}: {st['seconds']:.1f}s on {st['threads']} threads "
              f"({time.time() - t0:.0f}s total)", flush=True)
    ref = (acc / args.chunks).astype(np.float16 if args.res1024 else np.float32)
    np.savez_compressed(out, image=ref, spp=np.int64(spc * args.chunks),
                        frames=np.arange(1001, 1001 + args.chunks), max_depth=np.int64(8))
    print(f"wrote {out}: mean {ref.astype(np.float64).mean():.6f}", flush=True)


if __name__ == "__main__":
    main()
