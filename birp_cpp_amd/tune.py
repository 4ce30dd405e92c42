"""Tune the fast-path kernels (tile shape, lanes per ray, unroll) for the standard workloads and
write the table that batches, fast single renders and the server read:

    python -m birp_cpp_amd.tune [--out birp_cpp_amd/tuning.jsonl]

The views cover every (image-size class, samples class) the workloads use: the 256 orbit poses
of the batched workload (three image sizes, steps 0.1 and 0.25 RE), the 1600-sample views
of the precision workload, and 0.5 RE steps as the serving workload draws them. Each
class is tuned for FP32, BF16 and F16 storage on the 256^3 shell cube, by birp_autotune (every
candidate timed with HIP events on a batch of that class's views).
"""
from __future__ import annotations

import argparse
import os
import sys

from birp_cpp_amd import capi
from birp_cpp_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))


def tuning_views():
    views = W.orbit_views_mixed(256)
    views += W.orbit_views_576x336_n1600(32)
    pos = capi.synth_orbit(48, 9)
    views += [W.view(*W.SIZES[k % 3], 0.5, pos=tuple(float(c) for c in p), math=capi.MATH_FAST) for k, p in enumerate(pos)]
    return views


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.jsonl"))
    args = ap.parse_args(argv)
    dev = capi.Device(0)
    vals, org, spc, _ = W.cube_256()
    views = tuning_views()
    for dtype in (capi.DTYPE_F32, capi.DTYPE_BF16, capi.DTYPE_F16):
        vol = capi.Volume.upload(dev, vals, org, spc, dtype=dtype)
        try:
            print(f"tuning dtype {dtype} over {len(views)} views", file=sys.stderr, flush=True)
            capi.autotune(vol, views, args.out)
        finally:
            vol.free()
    dev.close()
    print(open(args.out).read(), end="")
    return 0


if __name__ == "__main__":
    sys.exit(main())
