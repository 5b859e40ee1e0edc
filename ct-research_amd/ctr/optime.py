"""Time each operator of the library on the current GPU.

    python -m ctr.optime [--views 360] [--n 256] [--reps 5] [--ops cone_fp,fdk_bp,...]

Runs on the 256^3 / 512^2 cone workload (1 mm voxels, SOD 600, SDD 1200,
0.5 mm isocentre pitch) or the 512^3 / 1024 x 768 one (--n 512), with the
3-D Shepp-Logan phantom forward projected on the device, and prints one line
per operator: device time per call (HIP events on the stream the library
launches on).  Use it under rocprofv3 to attribute kernel time per operator.

The par_* operators run the parallel-beam workload of FBP_parallel_GPU.m:
512 x 512 pixels over 60 mm, 360 one-degree views, t from -47 mm in steps
of 1.1 pixel (730 bins), on the 2-D Shepp-Logan phantom.  The curved_*
operators run the short scan of ImplementUIData.m: 512 x 512 x 50 voxels,
800 views over 240 degrees, 80 rows x 936 channels (uniform synthetic
channel angles, uniform random projections).  fan_sysmat builds the
system matrix of the fan configuration C1 (256^2 over 256 mm, 360 views, 512
equiangular bins at 1e-3 rad, SOD 500); fdk_error_slice runs
IterativeFDK.m's 128^3 / 360-view slice; cnn_forward runs the learned-update
network (4 layers, 32 channels, packaged weights) on the workload's volume.
"""
import argparse
import os
import sys
import time

OPS = ("cone_fp", "cone_filter", "fdk_bp", "cone_bp_ref", "cone_adjoint", "sirt3d",
       "par_fp", "par_filter", "par_fbp", "par_adjoint", "curved_filter", "curved_bp", "fan_sysmat",
       "fdk_error_slice", "cnn_forward")


def parallel_workload():
    """FBP_parallel_GPU.m:16-35: 512^2 over Size 60, t = -47 : 1.1 * 60/512 : 47, theta = 1..360 degrees."""
    import numpy as np
    from . import _lib
    res = 60.0 / 512
    t = np.arange(-47.0, 47.0 + 1e-9, 1.1 * res)
    theta = np.deg2rad(np.arange(1, 361))
    return _lib.ParGeom(512, 512, (60.0, 60.0), theta, t)


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
    if any(o.startswith("par_") for o in ops):
        gp = parallel_workload()
        img = torch.from_numpy(phantoms.shepp_logan_2d(512).astype("float32").ravel()).cuda()
        pproj = ctx.par_fp(gp, img)
        pfilt = ctx.par_filter(gp, pproj)
        pout = torch.empty_like(img)
        torch.cuda.synchronize()
        print(f"parallel workload: 512^2 image, {len(gp.theta)} views, {len(gp.t)} bins", flush=True)
    if any(o.startswith("curved_") for o in ops):
        from . import uidata
        gc = uidata.scan_geometry(uidata.synthetic_detector_angles())
        cproj = torch.rand(gc.nray, device="cuda")
        cfilt = ctx.curved_filter(gc, cproj)
        cvol = torch.empty(gc.nvox, device="cuda")
        torch.cuda.synchronize()
        print(f"curved workload: {gc.nt}x{gc.ns}x{gc.nz}, {len(gc.beta)} views, {len(gc.xi)}x{len(gc.gamma)} detector",
              flush=True)
    if "fan_sysmat" in ops:
        import numpy as np
        gam = 1e-3 * (np.arange(512) - 255.5)
        gf = ctr.FanGeom(256, 256, (256.0, 256.0), 500.0, np.deg2rad(np.arange(1, 361)), 500.0 * np.tan(gam))
    if "fdk_error_slice" in ops:
        import numpy as np
        from .geometry import iterative_fdk_args
        ea = iterative_fdk_args(128)
        ge = ctr.FdkErrGeom(128, 128, 128, ea["size"], ea["D"], ea["beta"], ea["u"], ea["p"], ea["g"], 63)
        evol = torch.rand(ge.nvox, device="cuda")
        eout = torch.empty(ge.nout, device="cuda")
    if "cnn_forward" in ops:
        import numpy as np
        from . import learned
        net = ctr.Cnn(ctx, ctr.cnn_init_params(0, 0.0))
        with np.load(learned.WEIGHTS, allow_pickle=False) as f:
            net.set_params(f["params"])
        cx = torch.rand(args.n ** 3, device="cuda")
        cg = torch.randn(args.n ** 3, device="cuda")
    calls = {
        "cone_fp": lambda: ctx.cone_fp(g, vol, out=proj),
        "cone_filter": lambda: ctx.cone_filter(g, proj, out=filt),
        "fdk_bp": lambda: ctx.fdk_bp(g, filt, out=out),
        "cone_bp_ref": lambda: ctx.cone_bp_ref(g, proj, out=out),
        "cone_adjoint": lambda: ctx.cone_bp_adjoint(g, proj, out=out),
        "par_fp": lambda: ctx.par_fp(gp, img, out=pproj),
        "par_filter": lambda: ctx.par_filter(gp, pproj, out=pfilt),
        "par_fbp": lambda: ctx.par_fbp(gp, pproj, out=pout),
        "par_adjoint": lambda: ctx.par_bp_adjoint(gp, pproj, out=pout),
        "curved_filter": lambda: ctx.curved_filter(gc, cproj, out=cfilt),
        "curved_bp": lambda: ctx.curved_bp(gc, cfilt, out=cvol),
        "fan_sysmat": lambda: ctx.fan_sysmat(gf, device=True),
        "fdk_error_slice": lambda: ctx.fdk_error_slice(ge, evol, out=eout),
        "cnn_forward": lambda: net.forward((args.n, args.n, args.n), cx, cg),
    }
    for o in ops:
        if o == "sirt3d":
            # a call costs a set-up (column sums V = A'1, the first residual) plus the
            # iterations: time 1 and 1 + n iterations and separate the two
            ctx.sirt3d(g, proj, 1)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.sirt3d(g, proj, 1)
            torch.cuda.synchronize()
            t1 = time.perf_counter() - t0
            t0 = time.perf_counter()
            ctx.sirt3d(g, proj, 1 + args.sirt_iters)
            torch.cuda.synchronize()
            tn = time.perf_counter() - t0
            it = (tn - t1) / args.sirt_iters
            print(f"{o:14s} {1e3 * it:9.3f} ms per iteration, set-up {1e3 * (t1 - it):.3f} ms per call "
                  f"(wall, {args.sirt_iters} iterations)", flush=True)
        else:
            reps = 1 if (o == "cone_adjoint" and args.n == 512) or o == "fdk_error_slice" else args.reps
            print(f"{o:14s} {timed(calls[o], reps):9.3f} ms", flush=True)
    ctx.close()


if __name__ == "__main__":
    sys.exit(main())
