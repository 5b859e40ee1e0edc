"""Every operator twice on the same input gives the same bits (no atomics in
any accumulation, fixed reduction orders, race-free filter passes)."""
import numpy as np
import pytest

import ctr
from ctr.geometry import iterative_fdk_args

pytestmark = pytest.mark.gpu


def _twice(f):
    a = np.array(f(), copy=True)
    b = np.array(f(), copy=True)
    return a, b


def test_operators_are_deterministic(ctx):
    rng = np.random.default_rng(41)
    g = ctr.geometry.cone_circular(32, 1.0, 24, 64, 48, 1.0, 150.0, 300.0)
    vol = rng.random(g.nvox).astype(np.float32)
    proj = rng.random(g.nray).astype(np.float32)
    cases = {
        "cone_fp": lambda: ctx.cone_fp(g, vol),
        "cone_filter": lambda: ctx.cone_filter(g, proj),
        "fdk_bp": lambda: ctx.fdk_bp(g, proj),
        "cone_bp_ref": lambda: ctx.cone_bp_ref(g, proj),
        "cone_adjoint": lambda: ctx.cone_bp_adjoint(g, proj),
        "sirt3d": lambda: ctx.sirt3d(g, proj, 3)[0],
    }
    gp = ctr.ParGeom(48, 48, (48.0, 48.0), np.deg2rad(np.arange(1, 61)), np.arange(-30, 30.5, 0.75))
    img = rng.random(gp.nvox).astype(np.float32)
    pp = rng.random(gp.nray).astype(np.float32)
    cases.update({
        "par_fp": lambda: ctx.par_fp(gp, img),
        "par_fbp": lambda: ctx.par_fbp(gp, pp),
        "par_adjoint": lambda: ctx.par_bp_adjoint(gp, pp),
    })
    a = iterative_fdk_args(16, n_views=24, pint=2.0)
    ge = ctr.FdkErrGeom(16, 16, 16, a["size"], a["D"], a["beta"], a["u"], a["p"], a["g"], 5)
    ve = rng.random(16 ** 3).astype(np.float32)
    cases["fdk_error_slice"] = lambda: ctx.fdk_error_slice(ge, ve)
    for name, f in cases.items():
        x, y = _twice(f)
        assert np.array_equal(x, y), name


def test_fp_masked_voxel_reads_are_bit_identical(ctx, monkeypatch):
    """The forward projector's masked second/third voxel reads (the default
    above 256 MiB of volume, forced here through CTR_FP_SKIP) add exactly the
    same terms as the unmasked walk: projections and SIRT iterates agree bit
    for bit."""
    rng = np.random.default_rng(43)
    g = ctr.geometry.cone_circular(40, 1.0, 30, 72, 56, 1.0, 150.0, 300.0)
    vol = rng.random(g.nvox).astype(np.float32)
    proj = rng.random(g.nray).astype(np.float32)
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("CTR_FP_SKIP", flag)
        c = ctr.Context(0)
        try:
            out[flag] = (np.array(c.cone_fp(g, vol), copy=True), np.array(c.sirt3d(g, proj, 3)[0], copy=True))
        finally:
            c.close()
    assert np.array_equal(out["0"][0], out["1"][0])
    assert np.array_equal(out["0"][1], out["1"][1])


def test_fp_non_finite_voxels_stay_on_their_rays(ctx, monkeypatch):
    """The library asks for finite volumes.  A non-finite voxel makes the
    projection of the rays through it non-finite in both read paths; a ray
    that only grazes it (a zero-length segment) may come out finite or not,
    depending on the path.  Every ray that passes nowhere near it keeps its
    value, bit for bit, in both paths."""
    rng = np.random.default_rng(44)
    g = ctr.geometry.cone_circular(40, 1.0, 30, 72, 56, 1.0, 150.0, 300.0)
    vol = rng.random((40, 40, 40)).astype(np.float32)
    k = (17, 22, 9)
    bad = vol.copy()
    bad[k] = np.inf
    near = np.zeros_like(vol)
    near[k[0] - 1:k[0] + 2, k[1] - 1:k[1] + 2, k[2] - 1:k[2] + 2] = 1.0
    one = np.zeros_like(vol)
    one[k] = 1.0
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("CTR_FP_SKIP", flag)
        c = ctr.Context(0)
        try:
            out[flag] = np.array(c.cone_fp(g, bad.ravel()), copy=True)
            ref = np.array(c.cone_fp(g, vol.ravel()), copy=True)
            through = np.array(c.cone_fp(g, one.ravel()), copy=True) > 0
            close = np.array(c.cone_fp(g, near.ravel()), copy=True) > 0
        finally:
            c.close()
        assert not np.isfinite(out[flag][through]).any(), flag
        assert np.array_equal(out[flag][~close], ref[~close]), flag
    assert through.sum() > 0 and (~close).sum() > 0.9 * g.nray
