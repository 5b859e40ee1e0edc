"""The allocation-free ramp filter (ctr_cone_filter_noalloc) through the C
ABI against ctr_cone_filter on the same device inputs: two calls in a row,
the second on a smaller geometry (its tables overwrite the first's in the
library's fixed device array), then a direct-convolution geometry and the
argument checks."""
import ctypes

import numpy as np
import pytest

import ctr
from ctr import geometry
from conftest import errs, report

pytestmark = pytest.mark.gpu


def _noalloc(ctx, g, filt, proj, out, flags=ctr.DEVICE_PTRS):
    fn = ctr.lib().ctr_cone_filter_noalloc
    return fn(ctx.h, ctypes.byref(g.c), int(filt), ctypes.c_void_p(proj), ctypes.c_void_p(out), int(flags))


@pytest.mark.parametrize("filt", [ctr.FILTER_SL_HALF, ctr.FILTER_RL_HALF])
def test_noalloc_filter_matches_filter(ctx, filt):
    import torch
    cases = [geometry.config_cone_256(360),                                  # 512 columns: FFT of 1024
             geometry.cone_circular(96, 1.0, 45, 200, 150, 1.2, 400.0, 800.0),   # smaller: FFT of 512
             geometry.cone_circular(64, 1.0, 12, 1100, 40, 0.2, 400.0, 800.0)]   # 1100 columns: direct sums
    gen = torch.Generator(device="cuda").manual_seed(71 + filt)
    for g in cases:
        proj = torch.rand(g.nray, device="cuda", generator=gen)
        ref = ctx.cone_filter(g, proj, filt)
        out = torch.full((g.nray,), float("nan"), device="cuda")
        ctx._bind_torch_stream()
        st = _noalloc(ctx, g, filt, proj.data_ptr(), out.data_ptr())
        assert st == ctr._lib.OK, ctr.lib().ctr_last_error(ctx.h).decode()
        torch.cuda.synchronize()
        e, m = errs(out.cpu().numpy(), ref.cpu().numpy())
        report(f"cone_filter_noalloc[{len(g.beta)}x{len(g.xi)}x{len(g.p)}, filter {filt}]", rel_l2=e, max_rel=m,
               identical=bool(torch.equal(out, ref)))
        assert e < 1e-6 and m < 1e-5


def test_noalloc_filter_rejects_what_it_cannot_hold(ctx):
    import torch
    g = geometry.cone_circular(32, 1.0, 4, 64, 48, 1.0, 200.0, 400.0)
    x = torch.zeros(g.nray, device="cuda")
    assert _noalloc(ctx, g, 0, x.data_ptr(), x.data_ptr(), flags=0) == ctr._lib.EINVAL      # host pointers
    assert b"device pointers" in ctr.lib().ctr_last_error(ctx.h)
    big = geometry.cone_circular(32, 1.0, 2, 2100, 8, 0.1, 200.0, 400.0)                  # > 2048 columns
    y = torch.zeros(big.nray, device="cuda")
    assert _noalloc(ctx, big, 0, y.data_ptr(), y.data_ptr()) == ctr._lib.ETOOBIG
    assert _noalloc(ctx, g, 7, x.data_ptr(), x.data_ptr()) == ctr._lib.EINVAL              # unknown filter
