"""Curved (cylindrical, equiangular) cone-beam short-scan FDK
(ImplementUI.cu) on the GPU against the FP64 oracle, and the full UI-scan
geometry against an analytic ball.  Tolerances per test; the HIP path
filters in FP64 (FFT, one spectrum per detector row) and backprojects in
FP32."""
import numpy as np
import pytest

from analytic import curved_ball_projections
from conftest import errs, report

pytestmark = pytest.mark.gpu


def curved_case(O, start=0.0):
    """48 x 44 x 8 voxels (anisotropic), 223 channels, 24 rows, a short scan
    of pi + fan with view and channel pitches that share no rational ratio
    (no Parker region boundary falls exactly on a sample)."""
    import ctr
    D, Dsd = 150.0, 300.0
    gamma = 0.00213 * (np.arange(223) - 111.3)
    xi = 1.07 * (np.arange(24) - 11.5)
    fan = abs(gamma[-1] - gamma[0])
    dbeta = 0.0101
    beta = start + dbeta * np.arange(int(np.ceil((np.pi + fan) / dbeta)) + 1)
    size = (48.0, 44 * 1.05, 8 * 0.9)
    args = (48, 44, 8, size, D, Dsd, beta, gamma, xi, start, fan)
    return ctr.CurvedGeom(*args), O.Curved(*args), (D, Dsd, beta, gamma, xi)


@pytest.mark.parametrize("start", [0.0, 0.027])
def test_curved_filter(ctx, oracle, start):
    g, og, _ = curved_case(oracle, start)
    y = np.random.default_rng(18).random(g.nray).astype(np.float32)
    e, m = errs(ctx.curved_filter(g, y), oracle.curved_filter(og, y))
    report(f"curved_filter[start={start}]", rel_l2=e, max_rel=m)
    assert e < 1e-6 and m < 1e-5


@pytest.mark.parametrize("quirks", [False, True])
def test_curved_bp(ctx, oracle, quirks):
    import ctr
    g, og, _ = curved_case(oracle, 0.027)
    y = np.random.default_rng(19).random(g.nray).astype(np.float32)
    gpu = ctx.curved_bp(g, y, flags=ctr.REF_QUIRKS if quirks else 0)
    with oracle.interp_mode(1):
        ref = oracle.curved_bp(og, y, quirks=quirks)
    e, m = errs(gpu, ref)
    report(f"curved_bp[quirks={quirks}]", rel_l2=e, max_rel=m)
    assert e < 2e-5 and m < 2e-4


def test_curved_bp_footprints_beyond_lds(ctx, oracle):
    """Channels 0.0005 rad and rows 0.3 mm apart: a tile's footprint spans
    several hundred channels and more rows than an LDS buffer holds, so the
    backprojection reads global memory for those tiles."""
    import ctr
    D, Dsd = 120.0, 240.0
    gamma = 0.0005 * (np.arange(701) - 350.2)
    xi = 0.3 * (np.arange(60) - 29.5)
    fan = abs(gamma[-1] - gamma[0])
    beta = 0.05 * np.arange(int(np.ceil((np.pi + fan) / 0.05)) + 1)
    args = (24, 20, 8, (24.0, 20.0, 8.0), D, Dsd, beta, gamma, xi, 0.0, fan)
    g, og = ctr.CurvedGeom(*args), oracle.Curved(*args)
    y = np.random.default_rng(23).random(g.nray).astype(np.float32)
    gpu = ctx.curved_bp(g, y)
    with oracle.interp_mode(1):
        ref = oracle.curved_bp(og, y)
    e, m = errs(gpu, ref)
    report("curved_bp[beyond LDS]", rel_l2=e, max_rel=m)
    assert e < 2e-5 and m < 2e-4


def test_curved_fdk_matches_oracle_and_accumulates(ctx, oracle):
    import ctr
    g, og, (D, Dsd, beta, gamma, xi) = curved_case(oracle, 0.027)
    proj = curved_ball_projections(beta, gamma, xi, D, Dsd, (3.0, -2.0, 0.5), 12.0, 1.0)
    vol = ctx.curved_fdk(g, proj)
    with oracle.interp_mode(1):
        ref = oracle.curved_bp(og, oracle.curved_filter(og, proj).astype(np.float32))
    e, m = errs(vol, ref)
    twice = ctx.curved_fdk(g, proj, out=vol.copy(), flags=ctr.ACCUMULATE)
    e2, _ = errs(twice, 2 * vol)
    report("curved_fdk", rel_l2=e, max_rel=m, accumulate_rel_l2=e2)
    assert e < 2e-5 and m < 2e-4 and e2 < 1e-7


def test_curved_fdk_ui_scan_ball():
    """The UI scan of ImplementUIData.m (512 x 512 x 50 over 512 x 512 x 23 mm,
    800 views over 240 degrees, 80 rows x 936 channels, SOD 570, SDD 1060.2):
    the FDK of the analytic projections of a ball of radius 100 mm is its
    amplitude in every slice."""
    import ctr
    from ctr import uidata
    g = uidata.scan_geometry(uidata.synthetic_detector_angles())
    A = 0.02
    centre = (20.0, -30.0, 0.0)
    proj = curved_ball_projections(g.beta, g.gamma, g.xi, g.distance, g.distance_s2d, centre, 100.0, A)
    ctx = ctr.Context(0)
    vol = ctx.curved_fdk(g, proj).reshape(50, 512, 512)
    ctx.close()
    c = np.arange(512) + 0.5 - 256
    r2 = (c[None, :] - centre[0]) ** 2 + (c[:, None] - centre[1]) ** 2
    inner, ring = r2 < 80.0 ** 2, (r2 > 120.0 ** 2) & (r2 < 180.0 ** 2)
    means = [float(vol[k][inner].mean() / A) for k in range(50)]
    rings = [float(vol[k][ring].mean() / A) for k in range(50)]
    stds = [float(vol[k][inner].std() / A) for k in range(50)]
    report("curved_fdk_ui_ball", inner_min=min(means), inner_max=max(means), inner_std_max=max(stds),
           ring_absmax=max(abs(v) for v in rings))
    # the FP64 oracle on a 64 x 64 x 10 grid of the same scan: means within 2e-4, rings 6e-4
    assert min(means) > 0.998 and max(means) < 1.002 and max(stds) < 2e-3
    assert max(abs(v) for v in rings) < 2e-3
