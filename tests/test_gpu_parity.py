"""GPU parity: every HIP operator against the FP64 CPU oracle on the same
seeded inputs, through the C ABI.  Tolerances are stated per test; the
HIP path computes in FP32 (FP64 per-ray setup for the projector, FP64
accumulation in the ramp filter), the oracle in FP64."""
import numpy as np
import pytest

from conftest import errs, report

pytestmark = pytest.mark.gpu


def cone_case(name):
    """Small cone geometries that exercise every kernel path: partial tiles,
    anisotropic voxels, truncated detectors, a detector so fine that tile
    footprints overflow the LDS buffer, descending sample arrays."""
    d = dict(nt=40, ns=36, nz=20, size=(40.0, 36 * 1.1, 20 * 0.9), D=120.0, nb=30, np_=64, nx=48,
             pp=1.0, px=0.8, desc=False)
    if name == "truncated":
        d.update(np_=40, nx=20, pp=0.7, px=0.7)
    elif name == "fine":
        d.update(nt=24, ns=20, nz=18, size=(24.0, 20.0, 18.0), np_=200, nx=150, pp=0.15, px=0.15, nb=12)
    elif name == "descending":
        d.update(desc=True)
    elif name == "exact":
        d.update(pp=0.5, px=0.5)
    beta = (2 * np.pi / d["nb"]) * np.arange(1, d["nb"] + 1)
    p = d["pp"] * (np.arange(d["np_"]) - (d["np_"] - 1) / 2)
    xi = d["px"] * (np.arange(d["nx"]) - (d["nx"] - 1) / 2) + (0.0 if name == "exact" else 0.05)
    if d["desc"]:
        p = p[::-1].copy()
        xi = xi[::-1].copy()
    return d, beta.astype(np.float32), p.astype(np.float32), xi.astype(np.float32)


def make_cone(name, O):
    import ctr
    d, beta, p, xi = cone_case(name)
    g = ctr.ConeGeom(d["nt"], d["ns"], d["nz"], d["size"], d["D"], beta, p, xi)
    og = O.Cone(d["nt"], d["ns"], d["nz"], d["size"], d["D"], beta, p, xi)
    return g, og


CASES = ["base", "truncated", "fine", "descending", "exact"]


@pytest.mark.parametrize("case", CASES)
def test_cone_fp(ctx, oracle, case):
    g, og = make_cone(case, oracle)
    print(f"[case] cone_fp {case} vol={g.vol_shape} proj={g.proj_shape}", flush=True)
    rng = np.random.default_rng(1)
    vol = rng.random(g.nvox).astype(np.float32)
    gpu = ctx.cone_fp(g, vol)
    ex = oracle.cone_fp_exact(og, vol)
    rf = oracle.cone_fp_ref(og, vol)
    e1, m1 = errs(gpu, ex)
    e2, m2 = errs(gpu, rf)
    report(f"cone_fp[{case}]", rel_l2_exact=e1, max_rel_exact=m1, rel_l2_ref=e2, max_rel_ref=m2)
    assert e1 < 2e-6 and m1 < 2e-5
    assert e2 < 2e-6 and m2 < 2e-5


def test_cone_fp_nonuniform_detector(ctx, oracle):
    """The forward projector reads per-sample positions (ProjectionCone_3D.cu:20-21),
    so an equiangular (p = D tan gamma) detector works for projection."""
    import ctr
    D = 150.0
    gam = 0.004 * (np.arange(70) - 34.5)
    p = (D * np.tan(gam)).astype(np.float32)
    xi = (0.9 * (np.arange(30) - 14.5)).astype(np.float32)
    beta = ((2 * np.pi / 17) * np.arange(1, 18)).astype(np.float32)
    g = ctr.ConeGeom(32, 32, 16, (32.0, 32.0, 16.0), D, beta, p, xi)
    og = oracle.Cone(32, 32, 16, (32.0, 32.0, 16.0), D, beta, p, xi)
    vol = np.random.default_rng(2).random(g.nvox).astype(np.float32)
    e, m = errs(ctx.cone_fp(g, vol), oracle.cone_fp_exact(og, vol))
    report("cone_fp[nonuniform]", rel_l2=e, max_rel=m)
    assert e < 2e-6 and m < 2e-5


@pytest.mark.parametrize("case", CASES)
def test_cone_bp_ref(ctx, oracle, case):
    g, og = make_cone(case, oracle)
    print(f"[case] cone_bp_ref {case}", flush=True)
    y = np.random.default_rng(3).random(g.nray).astype(np.float32)
    gpu = ctx.cone_bp_ref(g, y)
    with oracle.interp_mode(1):
        ref = oracle.cone_bp_ref(og, y)
    e, m = errs(gpu, ref)
    report(f"cone_bp_ref[{case}]", rel_l2=e, max_rel=m)
    assert e < 1e-5 and m < 1e-4
    if case == "exact":   # float-exact grid: identical to the reference's own grid
        e0, m0 = errs(gpu, oracle.cone_bp_ref(og, y))
        report("cone_bp_ref[exact,ref-grid]", rel_l2=e0, max_rel=m0)
        assert e0 < 1e-5 and m0 < 1e-4


@pytest.mark.parametrize("case", CASES)
def test_fdk_bp(ctx, oracle, case):
    g, og = make_cone(case, oracle)
    print(f"[case] fdk_bp {case}", flush=True)
    # zero-mean data like filtered projections: the per-view terms cancel, so the FP32
    # relative error of the sum is a few times that of the terms (bound 3e-5)
    y = np.random.default_rng(4).standard_normal(g.nray).astype(np.float32)
    gpu = ctx.fdk_bp(g, y)
    with oracle.interp_mode(1):
        ref = oracle.fdk_bp(og, y)
    e, m = errs(gpu, ref)
    report(f"fdk_bp[{case}]", rel_l2=e, max_rel=m)
    assert e < 3e-5 and m < 1e-4
    if case == "exact":
        e0, m0 = errs(gpu, oracle.fdk_bp(og, y))
        report("fdk_bp[exact,ref-grid]", rel_l2=e0, max_rel=m0)
        assert e0 < 3e-5 and m0 < 1e-4


def test_fdk_bp_ref_quirks(ctx, oracle):
    """CTR_REF_QUIRKS reproduces the stale sample above the detector (FDK.cu:100-123)."""
    import ctr
    g, og = make_cone("truncated", oracle)
    y = np.random.default_rng(5).random(g.nray).astype(np.float32)
    gpu = ctx.fdk_bp(g, y, flags=ctr.REF_QUIRKS)
    with oracle.interp_mode(1):
        ref_q = oracle.fdk_bp(og, y, quirks=True)
        ref_n = oracle.fdk_bp(og, y, quirks=False)
    e, m = errs(gpu, ref_q)
    report("fdk_bp[quirks]", rel_l2=e, max_rel=m, quirk_effect=errs(ref_q, ref_n)[0])
    assert errs(ref_q, ref_n)[0] > 1e-2   # the quirk is visible on this geometry
    assert e < 1e-5 and m < 2e-4


def test_fdk_accumulate(ctx, oracle):
    """View-chunked FDK (FDK_GPU.m:78-104): two half scans accumulated equal one full scan."""
    import ctr
    g, og = make_cone("base", oracle)
    y = np.random.default_rng(6).random(g.nray).astype(np.float32).reshape(g.proj_shape)
    full = ctx.fdk_bp(g, y)
    nb = len(g.beta)
    # same dbeta for both chunks: each chunk keeps two consecutive views first
    g1 = ctr.ConeGeom(g.nt, g.ns, g.nz, g.size, g.distance, g.beta[: nb // 2], g.p, g.xi)
    g2 = ctr.ConeGeom(g.nt, g.ns, g.nz, g.size, g.distance, g.beta[nb // 2:], g.p, g.xi)
    out = ctx.fdk_bp(g1, y[: nb // 2])
    ctx.fdk_bp(g2, y[nb // 2:], out=out, flags=ctr.ACCUMULATE)
    e, m = errs(out, full)
    report("fdk_bp[accumulate]", rel_l2=e, max_rel=m)
    assert e < 1e-6


@pytest.mark.parametrize("filt", [0, 1])
def test_cone_filter(ctx, oracle, filt):
    g, og = make_cone("base", oracle)
    print(f"[case] cone_filter filt={filt}", flush=True)
    y = np.random.default_rng(7).random(g.nray).astype(np.float32)
    gpu = ctx.cone_filter(g, y, filt=filt)
    ref = oracle.cone_filter(og, y, filt=filt)
    e, m = errs(gpu, ref)
    report(f"cone_filter[{filt}]", rel_l2=e, max_rel=m)
    assert e < 1e-6 and m < 1e-5


@pytest.mark.parametrize("np_", [33, 130, 512, 1000, 2048])
def test_filter_widths(ctx, oracle, np_):
    """Every template width of the row filter (K = 1..16, one and two rows per block)."""
    import ctr
    p = (0.37 * (np.arange(np_) - (np_ - 1) / 2)).astype(np.float32)
    xi = np.array([-0.5, 0.5], np.float32)
    beta = np.array([0.1, 0.2, 0.3], np.float32)
    g = ctr.ConeGeom(8, 8, 4, (8.0, 8.0, 4.0), 1000.0, beta, p, xi)
    og = oracle.Cone(8, 8, 4, (8.0, 8.0, 4.0), 1000.0, beta, p, xi)
    y = np.random.default_rng(np_).random(g.nray).astype(np.float32)
    for filt in (0, 1):
        e, m = errs(ctx.cone_filter(g, y, filt=filt), oracle.cone_filter(og, y, filt=filt))
        report(f"filter[np={np_},f={filt}]", rel_l2=e, max_rel=m)
        assert e < 1e-6 and m < 1e-5


@pytest.mark.parametrize("case", CASES)
def test_cone_adjoint(ctx, oracle, case):
    g, og = make_cone(case, oracle)
    print(f"[case] cone_adjoint {case}", flush=True)
    rng = np.random.default_rng(8)
    y = rng.random(g.nray).astype(np.float32)
    gpu = ctx.cone_bp_adjoint(g, y)
    ref = oracle.cone_adjoint(og, y)
    e, m = errs(gpu, ref)
    # adjoint identity with the GPU pair itself
    x = rng.random(g.nvox).astype(np.float32)
    ax = ctx.cone_fp(g, x).astype(np.float64)
    lhs = np.dot(ax, y.astype(np.float64))
    rhs = np.dot(x.astype(np.float64), gpu.astype(np.float64))
    adj = abs(lhs - rhs) / (np.linalg.norm(ax) * np.linalg.norm(y))
    report(f"cone_adjoint[{case}]", rel_l2=e, max_rel=m, adjoint_test=adj)
    assert adj < 1e-5
    assert e < 1e-4 and m < 1e-3


def test_cone_adjoint_cases_cover_both_plan_variants(ctx, oracle):
    """The exact adjoint launches a leaner variant when every footprint of
    the plan fits LDS, and the one with the global-memory path otherwise:
    the cases above run both."""
    glob = {case: ctx.bp_plan_stats(make_cone(case, oracle)[0], adjoint=True)["global"] for case in CASES}
    print(f"[case] adjoint footprints on the global path: {glob}", flush=True)
    assert any(v == 0 for v in glob.values()) and any(v > 0 for v in glob.values()), glob


def fan_case(equiangular=False, n=48, nb=40, nbins=80):
    D = 100.0
    size = (n * 0.9, n * 0.9)
    beta = (2 * np.pi / nb) * np.arange(1, nb + 1)
    if equiangular:
        gam = 0.008 * (np.arange(nbins) - (nbins - 1) / 2)
        p = D * np.tan(gam)
    else:
        p = 0.7 * (np.arange(nbins) - (nbins - 1) / 2)
    return n, size, D, beta, p


def make_fan(O, equiangular=False, single=False):
    import ctr
    n, size, D, beta, p = fan_case(equiangular)
    if single:
        beta = beta.astype(np.float32).astype(np.float64)
        p = p.astype(np.float32).astype(np.float64)
    return ctr.FanGeom(n, n, size, D, beta, p), O.Fan(n, n, size, D, beta, p)


def test_fan_fp(ctx, oracle):
    for eq in (False, True):
        g, og = make_fan(oracle, eq)
        img = np.random.default_rng(9).random(g.nvox).astype(np.float32)
        gpu = ctx.fan_fp(g, img)
        e1, m1 = errs(gpu, oracle.fan_fp_exact(og, img))
        e2, m2 = errs(gpu, oracle.fan_fp_ref(og, img))
        report(f"fan_fp[eq={eq}]", rel_l2_exact=e1, max_rel_exact=m1, rel_l2_ref=e2, max_rel_ref=m2)
        assert e1 < 2e-6 and m1 < 2e-5 and e2 < 2e-6 and m2 < 2e-5


@pytest.mark.parametrize("quirks", [False, True])
def test_fan_bp_ref(ctx, oracle, quirks):
    import ctr
    g, og = make_fan(oracle, single=True)
    y = np.random.default_rng(10).random(g.nray).astype(np.float32)
    gpu = ctx.fan_bp_ref(g, y, flags=ctr.REF_QUIRKS if quirks else 0)
    with oracle.interp_mode(1):
        ref = oracle.fan_bp_ref(og, y, quirks=quirks)
    e, m = errs(gpu, ref)
    report(f"fan_bp_ref[quirks={quirks}]", rel_l2=e, max_rel=m)
    assert e < 1e-5 and m < 2e-4


def test_fan_adjoint(ctx, oracle):
    g, og = make_fan(oracle)
    rng = np.random.default_rng(11)
    y = rng.random(g.nray).astype(np.float32)
    gpu = ctx.fan_bp_adjoint(g, y)
    e, m = errs(gpu, oracle.fan_adjoint(og, y))
    x = rng.random(g.nvox).astype(np.float32)
    ax = ctx.fan_fp(g, x).astype(np.float64)
    adj = abs(np.dot(ax, y) - np.dot(x.astype(np.float64), gpu)) / (np.linalg.norm(ax) * np.linalg.norm(y))
    report("fan_adjoint", rel_l2=e, max_rel=m, adjoint_test=adj)
    assert adj < 1e-5 and e < 1e-4 and m < 1e-3


@pytest.mark.parametrize("eq", [False, True])
@pytest.mark.parametrize("filt", [0, 1])
def test_fan_fbp(ctx, oracle, eq, filt):
    import ctr
    g, og = make_fan(oracle, eq, single=not eq)
    y = np.random.default_rng(12).random(g.nray).astype(np.float32)
    det = ctr.FAN_EQUIANGULAR if eq else ctr.FAN_FLAT
    f_gpu = ctx.fan_filter(g, y, detector=det, filt=filt)
    f_ref = oracle.fan_filter(og, y, equiangular=eq, filt=filt)
    e1, m1 = errs(f_gpu, f_ref)
    img = ctx.fan_fbp(g, y, detector=det, filt=filt)
    with oracle.interp_mode(1):
        ref = oracle.fan_fbp_bp(og, f_ref.astype(np.float32), equiangular=eq)
    e2, m2 = errs(img, ref)
    report(f"fan_fbp[eq={eq},f={filt}]", filt_rel_l2=e1, filt_max_rel=m1, fbp_rel_l2=e2, fbp_max_rel=m2)
    assert e1 < 1e-6 and m1 < 1e-5
    assert e2 < 2e-5 and m2 < 2e-4


def test_sirt3d(ctx, oracle):
    import ctr
    d, beta, p, xi = cone_case("base")
    g = ctr.ConeGeom(24, 24, 12, (24.0, 24.0, 12.0), 120.0, beta[::2], p, xi)
    og = oracle.Cone(24, 24, 12, (24.0, 24.0, 12.0), 120.0, beta[::2], p, xi)
    vol = np.zeros((12, 24, 24), np.float32)
    vol[3:9, 6:18, 5:19] = 1.0
    vol[5:7, 10:14, 9:13] = 0.5
    b = oracle.cone_fp_exact(og, vol.ravel()).astype(np.float32)
    x, r2, r1 = ctx.sirt3d(g, b, 20, lam=1.0, lo=0.0)
    xo, r2o, r1o = oracle.sirt3d(og, b, 20, lam=1.0, lo=0.0)
    e, m = errs(x, xo)
    report("sirt3d", rel_l2=e, max_rel=m, res_gpu=float(r2[-1]), res_oracle=float(r2o[-1]))
    assert e < 1e-4
    assert np.allclose(r2, r2o, rtol=1e-3, atol=1e-7) and np.allclose(r1, r1o, rtol=1e-3, atol=1e-9)
    assert np.all(np.diff(r2) <= 1e-7)   # monotone residual with the exact adjoint, lambda in (0, 2)
    assert x.min() >= 0.0


def test_sirt3d_normalisation_reuse():
    """V = 1/(A'1 + 1e-9) is kept per geometry between calls: a repeated call,
    a call after another geometry and the SIRT step all match a fresh context."""
    import ctr
    d, beta, p, xi = cone_case("base")
    g1 = ctr.ConeGeom(24, 24, 12, (24.0, 24.0, 12.0), 120.0, beta[::2], p, xi)
    g2 = ctr.ConeGeom(20, 28, 10, (20.0, 28.0, 10.0), 110.0, beta[1::2], p, xi)
    rng = np.random.default_rng(31)
    b1 = rng.random(g1.nray).astype(np.float32)
    b2 = rng.random(g2.nray).astype(np.float32)
    x0 = rng.random(g1.nvox).astype(np.float32)
    fresh = ctr.Context(0)
    ref1, _, _ = fresh.sirt3d(g1, b1, 4)
    ref2, _, _ = fresh.sirt3d(g2, b2, 4)
    fresh.close()
    fresh = ctr.Context(0)
    ref_step = fresh.sirt3d_step(g1, b1, x0)
    fresh.close()
    c = ctr.Context(0)
    try:
        a, _, _ = c.sirt3d(g1, b1, 4)
        again, _, _ = c.sirt3d(g1, b1, 4)
        other, _, _ = c.sirt3d(g2, b2, 4)
        back, _, _ = c.sirt3d(g1, b1, 4)
        step = c.sirt3d_step(g1, b1, x0)
    finally:
        c.close()
    for got, ref in ((a, ref1), (again, ref1), (other, ref2), (back, ref1), (step, ref_step)):
        assert np.array_equal(got, ref)


def test_sirt2d(ctx, oracle):
    g, og = make_fan(oracle)
    img = np.zeros((g.ns, g.nt), np.float32)
    img[10:38, 12:36] = 0.8
    img[20:26, 18:30] = 0.3
    b = oracle.fan_fp_exact(og, img.ravel()).astype(np.float32)
    x, r2, r1 = ctx.sirt2d(g, b, 30, lam=2.0, lo=0.0, hi=1.0)
    xo, r2o, r1o = oracle.sirt2d(og, b, 30, lam=2.0, lo=0.0, hi=1.0)
    e, m = errs(x, xo)
    report("sirt2d", rel_l2=e, max_rel=m, res_gpu=float(r2[-1]), res_oracle=float(r2o[-1]))
    assert e < 1e-4
    assert np.allclose(r2, r2o, rtol=1e-3, atol=1e-7)
    assert x.min() >= 0.0 and x.max() <= 1.0


def par_case(desc=False, n_t=90, nb=48, single=False):
    """48 x 40 anisotropic pixels; angles over 360 degrees including the axis
    directions; t offset by 0.013 so that no ray runs along a pixel edge;
    rays beyond the image (|t| up to 1.1 x the half diagonal)."""
    theta = np.concatenate([[0.0, np.pi / 2, np.pi, 1.5 * np.pi, np.pi / 4],
                            (2 * np.pi / (nb - 5)) * np.arange(1, nb - 4) + 0.01])
    t = 0.75 * (np.arange(n_t) - (n_t - 1) / 2) + 0.013
    if desc:
        t = t[::-1].copy()
    if single:
        theta = theta.astype(np.float32).astype(np.float64)
        t = t.astype(np.float32).astype(np.float64)
    return 48, 40, (48 * 0.8, 40 * 0.9), theta, t


def make_par(O, **kw):
    import ctr
    nt, ns, size, theta, t = par_case(**kw)
    return ctr.ParGeom(nt, ns, size, theta, t), O.Par(nt, ns, size, theta, t)


def test_fp_axis_and_diagonal_views(ctx, oracle):
    """Views along the axes and on the diagonals, exactly and within 1e-6 rad.
    Along an axis the central ray has no crossings of the other in-plane axis
    (zero direction component); on a diagonal the two in-plane steps are
    equal, so wavefronts there take the one-crossing walk while their
    neighbours walk slabs in lockstep.  A narrow fan (0.25 mm pitch at 300 mm)
    keeps whole wavefronts on one side of the diagonal test."""
    import ctr
    base = np.pi / 4 * np.arange(8)
    beta = np.concatenate([base, base + 1e-6, base - 1e-6]).astype(np.float32)
    p = (0.25 * (np.arange(128) - 63.5)).astype(np.float32)
    xi = (0.5 * (np.arange(24) - 11.5)).astype(np.float32)
    size = (32.0, 32.0, 16.0)
    g = ctr.ConeGeom(32, 32, 16, size, 300.0, beta, p, xi)
    og = oracle.Cone(32, 32, 16, size, 300.0, beta, p, xi)
    vol = np.random.default_rng(21).random(g.nvox).astype(np.float32)
    e, m = errs(ctx.cone_fp(g, vol), oracle.cone_fp_exact(og, vol))
    # parallel beam at theta = k pi/4.  t is offset from the pixel grid: a ray
    # lying exactly on a pixel boundary belongs to one side or the other by
    # the last bit of cos/sin (float on the device, double in the oracle)
    theta = (np.pi / 4 * np.arange(8)).astype(np.float64)
    t = ((np.arange(-20, 21) + 0.3) * 0.5).astype(np.float64)
    gp = ctr.ParGeom(24, 24, (24.0, 24.0), theta, t)
    ogp = oracle.Par(24, 24, (24.0, 24.0), theta, t)
    img = np.random.default_rng(22).random(gp.nvox).astype(np.float32)
    ep, mp = errs(ctx.par_fp(gp, img), oracle.par_fp_exact(ogp, img))
    report("fp[axis+diagonal]", cone_rel_l2=e, cone_max_rel=m, par_rel_l2=ep, par_max_rel=mp)
    assert e < 2e-6 and m < 2e-5
    assert ep < 2e-6 and mp < 2e-5


@pytest.mark.parametrize("desc", [False, True])
def test_par_fp(ctx, oracle, desc):
    g, og = make_par(oracle, desc=desc)
    img = np.random.default_rng(13).random(g.nvox).astype(np.float32)
    gpu = ctx.par_fp(g, img)
    e, m = errs(gpu, oracle.par_fp_exact(og, img))
    ones = ctx.par_fp(g, np.ones(g.nvox, np.float32))
    e1, m1 = errs(ones, oracle.par_fp_exact(og, np.ones(g.nvox, np.float32)))
    report(f"par_fp[desc={desc}]", rel_l2=e, max_rel=m, ones_rel_l2=e1)
    assert e < 2e-6 and m < 2e-5 and e1 < 2e-6


def test_par_adjoint(ctx, oracle):
    g, og = make_par(oracle)
    rng = np.random.default_rng(14)
    y = rng.random(g.nray).astype(np.float32)
    gpu = ctx.par_bp_adjoint(g, y)
    e, m = errs(gpu, oracle.par_adjoint(og, y))
    x = rng.random(g.nvox).astype(np.float32)
    ax = ctx.par_fp(g, x).astype(np.float64)
    adj = abs(np.dot(ax, y) - np.dot(x.astype(np.float64), gpu)) / (np.linalg.norm(ax) * np.linalg.norm(y))
    report("par_adjoint", rel_l2=e, max_rel=m, adjoint_test=adj)
    assert adj < 1e-5 and e < 1e-4 and m < 1e-3


@pytest.mark.parametrize("filt", [0, 1])
@pytest.mark.parametrize("desc", [False, True])
def test_par_fbp(ctx, oracle, filt, desc):
    g, og = make_par(oracle, desc=desc, single=True)
    y = np.random.default_rng(15).random(g.nray).astype(np.float32)
    f_gpu = ctx.par_filter(g, y, filt=filt)
    f_ref = oracle.par_filter(og, y, filt=filt)
    e1, m1 = errs(f_gpu, f_ref)
    img = ctx.par_fbp(g, y, filt=filt)
    with oracle.interp_mode(1):
        ref = oracle.par_fbp_bp(og, f_ref.astype(np.float32))
    e2, m2 = errs(img, ref)
    report(f"par_fbp[f={filt},desc={desc}]", filt_rel_l2=e1, filt_max_rel=m1, fbp_rel_l2=e2, fbp_max_rel=m2)
    assert e1 < 1e-6 and m1 < 1e-5
    assert e2 < 2e-5 and m2 < 2e-4


def test_par_filter_direct_path(ctx, oracle):
    """More than 1024 bins: the direct convolution instead of the FFT."""
    import ctr
    theta = np.deg2rad(np.arange(1, 7)).astype(np.float32).astype(np.float64)
    t = (0.05 * (np.arange(1500) - 749.5)).astype(np.float32).astype(np.float64)
    g, og = ctr.ParGeom(32, 32, (32.0, 32.0), theta, t), oracle.Par(32, 32, (32.0, 32.0), theta, t)
    y = np.random.default_rng(16).random(g.nray).astype(np.float32)
    e, m = errs(ctx.par_filter(g, y), oracle.par_filter(og, y))
    report("par_filter[1500 bins]", rel_l2=e, max_rel=m)
    assert e < 1e-6 and m < 1e-5


def test_par_fbp_disk_full_size(ctx):
    """512^2 over 360 one-degree views of 1024 bins: the FBP of the analytic
    projections of a uniform disk (2 A sqrt(R^2 - t^2)) is A inside it.
    Outside, one-degree views alias into streaks of up to ~0.13 A (the FP64
    oracle shows the same), which cancel on average over an annulus."""
    import ctr
    n, R, A = 512, 160.0, 0.8
    theta = np.deg2rad(np.arange(1, 361))
    t = 0.5 * (np.arange(1024) - 511.5)
    g = ctr.ParGeom(n, n, (float(n), float(n)), theta, t)
    proj = np.tile(2 * A * np.sqrt(np.maximum(R * R - t * t, 0)), len(theta)).astype(np.float32)
    img = ctx.par_fbp(g, proj).reshape(n, n)
    c = (np.arange(n) + 0.5) - n / 2
    rr = c[None, :] ** 2 + c[:, None] ** 2
    inner, ring = rr < (0.8 * R) ** 2, (rr > (1.2 * R) ** 2) & (rr < (1.5 * R) ** 2)
    report("par_fbp_disk[512]", inner_mean=float(img[inner].mean() / A), inner_std=float(img[inner].std() / A),
           ring_mean=float(img[ring].mean() / A))
    assert abs(img[inner].mean() / A - 1) < 1e-3 and img[inner].std() / A < 1e-3
    assert abs(img[ring].mean() / A) < 2e-3


def test_fdk_rows_in_path_matches_masked_path_and_oracle(ctx, oracle):
    """A detector narrow in u and tall in v: tiles whose footprint leaves the
    detector only in u take the clamped interior loop (rows-in).  The plan
    must contain such pairs; the result must equal the masked loop's
    (CTR_BP_ROWSIN=0) bit for bit and the oracle away from the u edges."""
    import os
    import ctr
    n = 48
    beta = ((2 * np.pi / 40) * np.arange(1, 41)).astype(np.float32)
    p = (0.9 * (np.arange(30) - 14.5)).astype(np.float32)          # +-13 mm: the 48 mm volume leaves it in u
    xi = (0.8 * (np.arange(110) - 54.5) + 0.05).astype(np.float32)  # +-43.6 mm: every row inside
    g = ctr.ConeGeom(n, n, n, (48.0, 48.0, 48.0), 150.0, beta, p, xi)
    og = oracle.Cone(n, n, n, (48.0, 48.0, 48.0), 150.0, beta, p, xi)
    stats = ctx.bp_plan_stats(g)
    y = np.random.default_rng(61).standard_normal(g.nray).astype(np.float32)
    fast = ctx.fdk_bp(g, y)
    os.environ["CTR_BP_ROWSIN"] = "0"
    try:
        masked = ctx.fdk_bp(g, y)
    finally:
        del os.environ["CTR_BP_ROWSIN"]
    with oracle.interp_mode(1):
        ref = oracle.fdk_bp(og, y)
    # columns whose centre projects within 1e-3 mm of the first/last column in some view
    c = (np.arange(n) + 0.5) - n / 2
    X, Y = c[None, :], c[:, None]
    near = np.zeros((n, n), bool)
    for b in beta.astype(np.float64):
        P = (X * np.cos(b) + Y * np.sin(b)) * 150.0 / (150.0 - (-X * np.sin(b) + Y * np.cos(b)))
        near |= (np.abs(P - p[0]) < 1e-3) | (np.abs(P - p[-1]) < 1e-3)
    keep = ~np.broadcast_to(near, (n, n, n)).ravel()
    e, m = errs(fast[keep], ref[keep])
    report("fdk_bp[rows-in]", **stats, identical_to_masked=bool(np.array_equal(fast, masked)), rel_l2=e, max_rel=m,
           edge_columns=int(near.sum()))
    assert stats["rows_in"] > 0.2 * stats["pairs"]
    assert np.array_equal(fast, masked)
    assert e < 3e-5 and m < 1e-4
