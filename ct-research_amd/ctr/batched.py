"""Batched plain SIRT over uneven cone-beam problems: throughput in
projector-voxel-updates per second, batched launches against one
ctr_sirt3d call per problem.

    python -m ctr.batched [--problems 16 --iters 50]

Problem i (seed 20261202 + i, xoshiro256**): N in {128, 192, 256, 320},
views in {90, 180, 360}, D x D detector with D uniform in 384..640, 1 mm
voxels, SOD 600, SDD 1200, isocentre pitch 1.5 N / D mm (the detector
about covers the volume diagonal), random-ellipsoid phantom (10-40
ellipsoids).  Throughput = sum_i 2 Nvox_i Nviews_i iters / time (one forward
and one backprojection per iteration, each counted as Nvox * Nviews).
"""
import argparse
import json
import sys
import time

import numpy as np

from . import _lib, geometry, phantoms

SEED0 = 20261202


def problem_spec(i):
    g = phantoms.Xoshiro256ss(SEED0 + i)
    n = (128, 192, 256, 320)[g.randint(0, 3)]
    views = (90, 180, 360)[g.randint(0, 2)]
    det = g.randint(384, 640)
    return n, views, det


def make_problems(ctx, count):
    import torch
    out = []
    for i in range(count):
        n, views, det = problem_spec(i)
        iso = 1.5 * n / det
        g = geometry.cone_circular(n, 1.0, views, det, det, iso * 2.0, 600.0, 1200.0)
        vol = phantoms.rasterize_3d_torch(phantoms.random_ellipsoids(SEED0 + i), n).reshape(-1).contiguous()
        b = ctx.cone_fp(g, vol)
        out.append((g, b, n, views, det))
    torch.cuda.synchronize()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", type=int, default=16)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--no-sequential", action="store_true")
    ap.add_argument("--autotune", action="store_true", help="tune every problem's geometry on first use")
    args = ap.parse_args(argv)
    import torch
    ctx = _lib.Context(0)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    probs = make_problems(ctx, args.problems)
    work = sum(2.0 * g.nvox * len(g.beta) * args.iters for g, *_ in probs)
    flags = _lib.AUTOTUNE if args.autotune else 0
    # warm-up: tables, footprints, code objects (and the tuning, when asked)
    ctx.sirt3d_batched([(g, b) for g, b, *_ in probs], 1, flags=flags)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xs, res = ctx.sirt3d_batched([(g, b) for g, b, *_ in probs], args.iters, flags=flags)
    torch.cuda.synchronize()
    t_batched = time.perf_counter() - t0
    out = {
        "metric": "SIRT throughput (batched, 16 uneven problems)",
        "unit": "Gvoxel-upd/s (2 Nvox Nviews per iteration)",
        "problems": [{"n": n, "views": v, "detector": d, "residual_final": round(float(r[-1]), 6)}
                     for (g, b, n, v, d), r in zip(probs, res)],
        "iters": args.iters, "autotune": bool(args.autotune),
        "batched_s": round(t_batched, 4), "batched_gups": round(work / t_batched / 1e9, 2),
    }
    if not args.no_sequential:
        for g, b, *_ in probs:
            ctx.sirt3d(g, b, 1, flags=flags)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for (g, b, *_), xb in zip(probs, xs):
            ctx.sirt3d(g, b, args.iters, residuals=False, flags=flags)
        torch.cuda.synchronize()
        t_seq = time.perf_counter() - t0
        out.update({"sequential_s": round(t_seq, 4), "sequential_gups": round(work / t_seq / 1e9, 2)})
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    sys.exit(main())
