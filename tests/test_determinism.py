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
