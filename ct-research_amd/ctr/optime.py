"""Time each operator of the library on the current GPU.

    python -m ctr.optime [--views 360] [--n 256] [--reps 5] [--ops cone_fp,fdk_bp,...]

Runs on the 256^3 / 512^2 cone workload (1 mm voxels, SOD 600, SDD 1200,
0.5 mm isocentre pitch) or the 512^3 / 1024 x 768 one (--n 512), with the
3-D Shepp-Logan phantom forward projected on the device, and prints one line
per operator: device time per call (HIP events on the stream the library
launches on).  Use it under rocprofv3 to attribute kernel time per operator.
"""
import argparse
import os
import sys
import time

OPS = ("cone_fp", "cone_filter", "fdk_bp", "cone_bp_ref", "cone_adjoint", "sirt3d")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--views", type=int, default=None, help="default 360 (n=256) or 720 (n=512)")
    ap.add_argument("--n", type=int, default=256, choices=(256, 512))
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--sirt-iters", type=int, default=5)
    ap.add_argument("--ops", default=",".join(OPS))
    args = ap.parse_args(argv)
    ops = [o for o in args.ops.split(",") if o]
    for o in ops:
        if o not in OPS:
            raise SystemExit(f"unknown op {o}; choose from {','.join(OPS)}")

    import torch
    import ctr
    from ctr import geometry, phantoms

    ctx = ctr.Context(0)
    st = torch.cuda.current_stream()
    ctx.set_stream(st.cuda_stream)
    if args.n == 256:
        g = geometry.config_cone_256(args.views or 360)
    else:
        g = geometry.cone_circular(512, 0.5, args.views or 720, 1024, 768, 0.8, 700.0, 1100.0)
    vol = torch.from_numpy(phantoms.phantom3d(args.n, "shepp-logan", 12).ravel()).cuda()
    proj = ctx.cone_fp(g, vol)
    filt = ctx.cone_filter(g, proj)
    out = torch.empty_like(vol)
    torch.cuda.synchronize()

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        for _ in range(reps):
            fn()
        b.record(st)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    print(f"workload: {args.n}^3 volume, {len(g.beta)} views, {len(g.p)}x{len(g.xi)} detector", flush=True)
    calls = {
        "cone_fp": lambda: ctx.cone_fp(g, vol, out=proj),
        "cone_filter": lambda: ctx.cone_filter(g, proj, out=filt),
        "fdk_bp": lambda: ctx.fdk_bp(g, filt, out=out),
        "cone_bp_ref": lambda: ctx.cone_bp_ref(g, proj, out=out),
        "cone_adjoint": lambda: ctx.cone_bp_adjoint(g, proj, out=out),
    }
    for o in ops:
        if o == "sirt3d":
            ctx.sirt3d(g, proj, 1)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.sirt3d(g, proj, args.sirt_iters)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / args.sirt_iters
            print(f"{o:14s} {ms:9.3f} ms per iteration (wall, {args.sirt_iters} iterations)", flush=True)
        else:
            reps = 1 if o == "cone_adjoint" and args.n == 512 else args.reps
            print(f"{o:14s} {timed(calls[o], reps):9.3f} ms", flush=True)
    ctx.close()


if __name__ == "__main__":
    sys.exit(main())
