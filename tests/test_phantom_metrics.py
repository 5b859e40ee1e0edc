"""Device phantom rasterizer and metrics against the host definitions
(phantom3d.m via ctr.phantoms.rasterize_3d, rRMSE.m, LocalError.m)."""
import numpy as np
import pytest

from conftest import report

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,n", [("shepp-logan", 64), ("modified shepp-logan", 65)])
def test_phantom_matches_host_rasterizer(ctx, kind, n):
    import ctr
    e = ctr.phantoms.ellipsoids_3d(kind)
    gpu = ctx.phantom_ellipsoids(e, n)
    host = ctr.phantoms.rasterize_3d(e, n)
    # same FP64 expression; cos/sin of phi may differ in the last bit between libraries,
    # which can move a voxel centre lying on an ellipsoid surface
    diff = np.abs(gpu - host)
    report(f"phantom[{kind},{n}]", voxels_differing=int((diff > 1e-6).sum()), max_diff=float(diff.max()))
    assert (diff > 1e-6).sum() <= 4


def test_phantom_random_ellipsoids_anisotropic_grid(ctx):
    import ctr
    e = ctr.phantoms.random_ellipsoids(20261202)
    gpu = ctx.phantom_ellipsoids(e[:, :8], 40, 36, 20, device=True).cpu().numpy()
    # host reference on the same anisotropic grid
    r = lambda n: (np.arange(n) - (n - 1) / 2.0) / ((n - 1) / 2.0)
    y, x, z = r(40)[None, None, :], r(36)[None, :, None], r(20)[:, None, None]
    ref = np.zeros((20, 36, 40))
    for A, a, b, c, x0, y0, z0, phi in e[:, :8]:
        ph = np.deg2rad(phi)
        cp, sp = np.cos(ph), np.sin(ph)
        sx, sy = cp * x0 + sp * y0, -sp * x0 + cp * y0
        u = (cp * x + sp * y - sx) ** 2 / a ** 2 + (-sp * x + cp * y - sy) ** 2 / b ** 2
        ref += A * ((u + (z - z0) ** 2 / c ** 2) <= 1.0)
    assert (np.abs(gpu - ref) > 1e-6).sum() <= 4


def test_metrics(ctx):
    import ctr
    rng = np.random.default_rng(34)
    ref = rng.random(100003).astype(np.float32)
    x = (ref + 0.01 * rng.standard_normal(ref.size)).astype(np.float32)
    r = ctx.rrmse(x, ref)
    le = ctx.local_error(x, ref)
    assert np.isclose(r, ctr.metrics.rrmse(x, ref), rtol=1e-12)
    assert np.isclose(le, ctr.metrics.local_error(x, ref), rtol=1e-12)
    import torch
    assert np.isclose(ctx.rrmse(torch.from_numpy(x).cuda(), torch.from_numpy(ref).cuda()), r, rtol=1e-12)
