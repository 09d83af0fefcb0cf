"""Writes tests/golden/cornell_256_fp64_65536spp.npz: the Cornell box
(256x256, max depth 8, NEE, RR 0.8) rendered by the CPU checker in FP64
(reference-image mode) with 65,536 samples per pixel, as 16 independent
frames of 4096 spp (frame indices 1001..1016) averaged in float64.

The image is the error reference for relMSE reports (the reference's
metric, src/render/passes/errormeasure/metrics.cu:80-90).  It is data
produced by this repository's own checker, not by the reference.

usage: python tests/golden/make_cornell_reference.py [--spp-per-chunk N] [--chunks K]
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--spp-per-chunk", type=int, default=4096)
    ap.add_argument("--chunks", type=int, default=16)
    ap.add_argument("--out", default=CORNELL_REFERENCE)
    args = ap.parse_args()
    import kiraray_amd as krr
    import oracle.binding as ob
    from conftest import CORNELL
    s = krr.Scene(CORNELL)
    acc = np.zeros((256, 256, 3), np.float64)
    t0 = time.time()
    for c in range(args.chunks):
        img, st = ob.render(s.desc, s.camera(0), 256, 256, 1001 + c, spp=args.spp_per_chunk, max_depth=8, nee=1, rr=0.8,
                            precision=64)
        acc += img[..., :3].astype(np.float64)
        print(f"chunk {c + 1}/{args.chunks}: {st['seconds']:.1f}s on {st['threads']} threads "
              f"({time.time() - t0:.0f}s total)", flush=True)
    ref = (acc / args.chunks).astype(np.float32)
    np.savez_compressed(args.out, image=ref, spp=np.int64(args.spp_per_chunk * args.chunks),
                        frames=np.arange(1001, 1001 + args.chunks), max_depth=np.int64(8))
    print(f"wrote {args.out}: mean {ref.mean():.6f}", flush=True)


if __name__ == "__main__":
    main()
