"""Siddon system-matrix export (ctr_*_sysmat) against the FP64 oracle's
per-ray segment lists, the projectors (A x equals the forward projection)
and the MATLAB builder gateways.  Segments shorter than 1e-7 mm (tie
slivers, where float rounding decides which side of a plane a crossing
falls) are dropped from both sides before the row comparison."""
import numpy as np
import pytest

from conftest import errs, report
from test_gpu_parity import cone_case, make_cone, make_fan, make_par

pytestmark = pytest.mark.gpu

SLIVER = 1e-7


def csr(rp, col, val, shape):
    import scipy.sparse as sp
    return sp.csr_matrix((val.astype(np.float64), col, rp), shape=shape)


def check_rows(rp, col, val, segments, rays):
    worst = 0.0
    for r in rays:
        c, v = col[rp[r]:rp[r + 1]], val[rp[r]:rp[r + 1]]
        oi, ol = segments(r)
        kg, ko = v > SLIVER, ol > SLIVER
        assert np.array_equal(c[kg], oi[ko]), r
        if kg.any():
            d = np.abs(v[kg] - ol[ko]) / np.maximum(ol[ko], 1.0)
            worst = max(worst, float(d.max()))
        assert len(np.unique(c)) == len(c) and np.all(v >= 0)
    return worst


def test_fan_sysmat(ctx, oracle):
    g, og = make_fan(oracle)
    rp, col, val = ctx.fan_sysmat(g)
    worst = check_rows(rp, col, val, lambda r: oracle.fan_segments(og, r), range(g.nray))
    x = np.random.default_rng(24).random(g.nvox).astype(np.float32)
    e, m = errs(csr(rp, col, val, (g.nray, g.nvox)) @ x.astype(np.float64), ctx.fan_fp(g, x))
    report("fan_sysmat", nnz=int(rp[-1]), worst_len_rel=worst, ax_vs_fp_rel_l2=e, ax_vs_fp_max_rel=m)
    assert worst < 1e-6 and e < 1e-6 and m < 1e-5


@pytest.mark.parametrize("desc", [False, True])
def test_par_sysmat(ctx, oracle, desc):
    g, og = make_par(oracle, desc=desc)
    rp, col, val = ctx.par_sysmat(g)
    worst = check_rows(rp, col, val, lambda r: oracle.par_segments(og, r), range(g.nray))
    x = np.random.default_rng(25).random(g.nvox).astype(np.float32)
    e, m = errs(csr(rp, col, val, (g.nray, g.nvox)) @ x.astype(np.float64), ctx.par_fp(g, x))
    report(f"par_sysmat[desc={desc}]", nnz=int(rp[-1]), worst_len_rel=worst, ax_vs_fp_rel_l2=e)
    assert worst < 1e-6 and e < 1e-6 and m < 1e-5


@pytest.mark.parametrize("case", ["base", "descending"])
def test_cone_sysmat(ctx, oracle, case):
    g, og = make_cone(case, oracle)
    rp, col, val = ctx.cone_sysmat(g)
    worst = check_rows(rp, col, val, lambda r: oracle.cone_segments(og, r), range(0, g.nray, 11))
    x = np.random.default_rng(26).random(g.nvox).astype(np.float32)
    A = csr(rp, col, val, (g.nray, g.nvox))
    e, m = errs(A @ x.astype(np.float64), ctx.cone_fp(g, x))
    # column sums are the exact adjoint of ones
    e2, _ = errs(np.asarray(A.sum(axis=0)).ravel(), ctx.cone_bp_adjoint(g, np.ones(g.nray, np.float32)))
    report(f"cone_sysmat[{case}]", nnz=int(rp[-1]), worst_len_rel=worst, ax_vs_fp_rel_l2=e, colsum_vs_adjoint=e2)
    assert worst < 1e-6 and e < 1e-6 and m < 1e-5 and e2 < 1e-5


def test_sysmat_device_pointers_and_capacity(ctx, oracle):
    import ctypes
    import ctr
    g, _ = make_fan(oracle)
    rp, col, val = ctx.fan_sysmat(g)
    drp, dcol, dval = ctx.fan_sysmat(g, device=True)
    assert np.array_equal(drp.cpu().numpy(), rp) and np.array_equal(dcol.cpu().numpy(), col)
    assert np.array_equal(dval.cpu().numpy(), val)
    # a capacity below the entry count is refused before anything is written
    nnz = ctypes.c_int64(int(rp[-1]) - 1)
    c2, v2 = np.zeros(int(rp[-1]), np.int32), np.zeros(int(rp[-1]), np.float32)
    st = ctr._lib.lib().ctr_fan_sysmat(ctx.h, ctypes.byref(g.c), rp.ctypes.data, c2.ctypes.data, v2.ctypes.data,
                                       ctypes.byref(nnz), 0)
    assert st == ctr._lib.EINVAL and not c2.any()
