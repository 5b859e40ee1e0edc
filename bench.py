"""Benchmark: FDK backprojection throughput (Gvoxel-updates/s) on one MI355X,
with the learned-SIRT time-to-residual comparison beside it.

Workload: 256^3 volume (1 mm voxels), 360 views over 360 deg, 512 x 512 flat
detector (1 mm pitch, SOD 600, SDD 1200 -> 0.5 mm at the isocentre).  The
3-D Shepp-Logan phantom is forward projected and ramp filtered on the GPU
before the timed region; a step is one FDK backprojection of all 360
filtered views into the volume, inputs resident in HBM.

The "m2" object times plain SIRT to its 200-iteration data residual against
the learned SIRT on the sparse-view problem (256^3, 90 views, Poisson noise
at I0 = 1e5; held-out random-ellipsoid phantom), see ctr/learned.py.

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ct-research_amd"))

HBM_PEAK_GBS = 8000.0
N_CU = 256
PEAK_CLOCK_GHZ = 2.4            # MI355X peak engine clock
FDK_PROFILE = os.path.join(ROOT, "profiles", "fdk_bp_256_360_sq_counters.txt")
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "fdk_bp_256_360_traffic.txt")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-m2", action="store_true", help="skip the learned-vs-plain SIRT comparison")
    ap.add_argument("--cpu-views", type=int, default=360, help="views in the CPU-baseline sample (all threads)")
    ap.add_argument("--cpu-views-1t", type=int, default=8, help="views in the 1-thread CPU-baseline sample")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="before the timed region, fill and then copy (elementwise multiply by 1) a buffer of this "
                         "many MiB: kernels of known traffic that calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def _profile_value(path, section, key):
    """A counter value from a committed rocprofv3 summary (`section:` block, `key value` lines)."""
    try:
        with open(path) as f:
            inside = False
            for line in f:
                if line.strip().endswith(":") and not line.startswith(" "):
                    inside = line.strip()[:-1] == section
                elif inside and line.split() and line.split()[0] == key:
                    return float(line.split()[1])
    except OSError:
        pass
    return None


def traffic_from_profiles():
    """HBM bytes per launch of the FDK kernel measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, calibrated with a copy kernel).  None when not measured."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(geom_args, n_views, n_views_1t):
    """The FP64 CPU restatement of the reference FDK backprojection
    (oracle/ctr_oracle.c or_fdk_bp) on bounded samples: the full 256^3 volume
    and the first n_views views with OpenMP over all host threads, and the
    first n_views_1t views on one thread."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    nt, ns, nz, size, D, beta, p, xi = geom_args

    def timed(views):
        og = O.Cone(nt, ns, nz, size, D, beta[:views], p, xi)
        y = np.random.default_rng(0).standard_normal(og.nray).astype(np.float32)
        t0 = time.perf_counter()
        O.fdk_bp(og, y)
        dt = time.perf_counter() - t0
        return nt * ns * nz * views / dt / 1e9, dt

    cores = O.num_threads()
    v_all, dt_all = timed(n_views)
    O.set_num_threads(1)
    try:
        v_1, dt_1 = timed(n_views_1t)
    finally:
        O.set_num_threads(cores)
    return {"value": v_all, "unit": "Gvoxel-upd/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(), "value_1thread": v_1,
            "sample": f"FP64 restatement of FDK.cu BackProjection, 256^3 volume x {n_views} of 360 views on "
                      f"{cores} threads ({dt_all:.1f} s, OpenMP); 1 thread: {n_views_1t} views ({dt_1:.1f} s)"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    import ctr
    from ctr import geometry, phantoms

    dev = torch.device("cuda:0")
    ctx = ctr.Context(0)
    stream = torch.cuda.current_stream()
    ctx.set_stream(stream.cuda_stream)

    g = geometry.config_cone_256(360)
    nvox, nviews, ndet = g.nvox, len(g.beta), len(g.p) * len(g.xi)
    vol = torch.from_numpy(phantoms.phantom3d(256, "shepp-logan", 12).ravel()).to(dev)
    proj = ctx.cone_fp(g, vol)
    filt = ctx.cone_filter(g, proj, filt=ctr.FILTER_SL_HALF)
    out = torch.empty(nvox, dtype=torch.float32, device=dev)
    if args.calib_copy_mb:
        # fill (writes n bytes) and an elementwise multiply (reads n, writes n):
        # plain kernels of known traffic, unlike copy_, which may use the DMA engine
        a = torch.empty(args.calib_copy_mb << 18, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.mul(a, 1.0, out=b)
        del a, b
    torch.cuda.synchronize()

    # filter timing (reported separately, not part of the metric)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ctx.cone_filter(g, proj, filt=ctr.FILTER_SL_HALF, out=filt)
    e0.record(stream)
    for _ in range(5):
        ctx.cone_filter(g, proj, filt=ctr.FILTER_SL_HALF, out=filt)
    e1.record(stream)
    torch.cuda.synchronize()
    filter_ms = e0.elapsed_time(e1) / 5

    for _ in range(args.warmup):
        ctx.fdk_bp(g, filt, out=out)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        ctx.fdk_bp(g, filt, out=out)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps     # one kernel per step on this stream
    ms_per_step = wall * 1000.0 / args.steps
    value = nvox * nviews / (ms_per_step * 1e-3) / 1e9

    # HBM roofline: algorithmic bytes of the streaming dataflow with K views per pass
    K = nviews
    alg_bytes = math.ceil(nviews / K) * 8 * nvox + 4 * nviews * ndet
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    # the bound that binds: VALU issue.  Wave-instructions per launch from the
    # committed counter profile of this kernel on this workload; the ceiling is
    # one wave64 VALU instruction per CU per clock at the peak engine clock.
    valu = _profile_value(FDK_PROFILE, "k_bp_tiled", "SQ_INSTS_VALU")
    valu_peak = N_CU * PEAK_CLOCK_GHZ * 1e9
    valu_roof = None
    if valu:
        valu_rate = valu / (kernel_ms * 1e-3)
        valu_roof = {"bound": "valu-issue", "achieved": round(valu_rate / 1e12, 4), "peak": round(valu_peak / 1e12, 4),
                     "unit": "T wave-instr/s", "frac": round(valu_rate / valu_peak, 4),
                     "valu_per_launch": valu, "valu_per_voxel_update": round(valu * 64 / (nvox * nviews), 3),
                     "lds_cap_gvupd": round(8 * N_CU * PEAK_CLOCK_GHZ, 1),
                     "source": "SQ_INSTS_VALU of k_bp_tiled in profiles/fdk_bp_256_360_sq_counters.txt"}
    res = {
        "metric": "Gvoxel-upd/s (FDK BP)",
        "value": round(value, 2),
        "unit": "Gvoxel-upd/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (3-D Shepp-Logan, 12 ellipsoids, forward projected and SL-half filtered on the GPU)",
        "config": {"workload": "fdk_bp 256^3 vol 1mm, 360 views, 512x512 flat det 1mm, SOD 600, SDD 1200",
                   "views_per_pass": K, "filter_ms": round(filter_ms, 4), "kernel_ms": round(kernel_ms, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic_from_profiles(),
                     "traffic_source": "FETCH_SIZE x 2 + WRITE_SIZE of k_bp_tiled, "
                                       "profiles/fdk_bp_256_360_traffic.txt",
                     # HBM is the contract's memory roofline but not what bounds this kernel
                     "binds": False, "binding_roofline": "valu_roofline"},
        "valu_roofline": valu_roof,
    }
    del proj, filt, out, vol
    torch.cuda.empty_cache()
    if not args.no_m2:
        from ctr import learned
        res["m2"] = learned.m2_measure(ctx)
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline((g.nt, g.ns, g.nz, g.size, g.distance, g.beta, g.p, g.xi),
                                           args.cpu_views, args.cpu_views_1t)
    print(json.dumps(res), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
