"""GPU SIRT-200 against the FP64 oracle's SIRT-200 on a reduced sparse-view
problem: the learned-SIRT comparison at 256^3 uses GPU SIRT-200 as its image
reference, because a 256^3 FP64 CPU run takes hours.

    python tests/sirt200_reference.py oracle OUT.npz [--n 96]   # CPU: problem + FP64 SIRT-200
    python tests/sirt200_reference.py gpu OUT.npz                # GPU: SIRT-200 on the same b; prints the comparison

Problem: n^3 random-ellipsoid phantom (seed 2000) over the 256 mm field of
view of the 256^3 case (voxel 256/n mm), 90 views, a 2n x 2n flat detector
at the same 1 mm physical pitch scaled to n (iso pitch 256/n mm at SOD 600,
SDD 1200), line integrals scaled to a maximum of 4, Poisson noise at
I0 = 1e5.  Not collected by pytest: the oracle half takes about 30 minutes
on 6 cores at n = 96.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path[:0] = [os.path.join(ROOT, "ct-research_amd"), os.path.join(ROOT, "oracle")]


def geometry_for(n):
    from ctr import geometry
    return geometry.cone_circular(n, 256.0 / n, 90, 2 * n, 2 * n, 2 * 256.0 / (2 * n), 600.0, 1200.0)


def oracle_half(out, n):
    import oracle as O
    from ctr import phantoms
    g = geometry_for(n)
    og = O.Cone(g.nt, g.ns, g.nz, g.size, g.distance, g.beta, g.p, g.xi)
    vol = phantoms.random_phantom(n, 2000).ravel()
    b0 = O.cone_fp_exact(og, vol)
    scale = 4.0 / b0.max()
    b = (phantoms.poisson_noise(b0 * scale, 1e5, 2000) / scale).astype(np.float32)
    t0 = time.perf_counter()
    x, r2, r1 = O.sirt3d(og, b, 200)
    np.savez(out, b=b, x=x, r2=r2, n=n, oracle_s=time.perf_counter() - t0, threads=O.num_threads())
    print(json.dumps(dict(n=n, residual_200=float(r2[-1]), seconds=time.perf_counter() - t0)), flush=True)


def gpu_half(path):
    import ctr
    with np.load(path, allow_pickle=False) as f:
        b, xo, r2o, n = f["b"], f["x"], f["r2"], int(f["n"])
    g = geometry_for(n)
    ctx = ctr.Context(0)
    try:
        x, r2, _ = ctx.sirt3d(g, b, 200)
    finally:
        ctx.close()
    e = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    m = np.abs(x - xo).max() / np.abs(xo).max()
    out = dict(n=n, image_rel_l2=float(e), image_max_rel=float(m), residual_gpu=float(r2[-1]),
               residual_oracle=float(r2o[-1]), residual_history_max_rel=float(np.max(np.abs(r2 - r2o) / r2o)),
               zero_voxels_gpu=float(np.mean(x == 0)), zero_voxels_oracle=float(np.mean(xo == 0)))
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("half", choices=("oracle", "gpu"))
    ap.add_argument("path")
    ap.add_argument("--n", type=int, default=96)
    a = ap.parse_args()
    if a.half == "oracle":
        oracle_half(a.path, a.n)
    else:
        gpu_half(a.path)
