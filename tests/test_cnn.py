"""The learned SIRT update: CNN kernels against a PyTorch fp32 reference of
the same network, and the residual-judged loop against plain SIRT.

The reference rounds inputs, weights and the ReLU activations to bfloat16
exactly where the kernels store BF16, and accumulates in fp32 (CPU), so the
remaining differences are FP32 summation order and BF16 rounding flips of
individual activations.  The gradient reference differentiates the same
graph with straight-through rounding; the kernels also round the deltas to
BF16, hence the looser gradient tolerance.
"""
import numpy as np
import pytest

import ctr
from ctr import geometry
from conftest import report

FWD_TOL = 5e-3     # rel-L2 of the network output
GRAD_TOL = 3e-2    # rel-L2 of each layer's weight gradient


def test_param_layout():
    p = ctr.cnn_init_params(0, 1.0)
    assert p.size == ctr.CNN_NUM_PARAMS == 57985
    layers = ctr.cnn_split_params(p)
    assert [w.shape for w, _ in layers] == [(32, 2, 3, 3, 3), (32, 32, 3, 3, 3), (32, 32, 3, 3, 3), (1, 32, 3, 3, 3)]
    assert [b.shape for _, b in layers] == [(32,), (32,), (32,), (1,)]
    assert np.all(ctr.cnn_init_params(0, 0.0)[-865:] == 0)   # last layer off: plain SIRT


def _torch_net(p, shape, x, g, gs, ste=False):
    import torch
    import torch.nn.functional as F
    nx, ny, nz = shape

    def bf(t):
        r = t.to(torch.bfloat16).float()
        return t + (r - t).detach() if ste else r

    inp = torch.stack([torch.from_numpy(np.asarray(x, np.float32).reshape(nz, ny, nx)),
                       torch.from_numpy(np.asarray(g, np.float32).reshape(nz, ny, nx)) * gs])[None]
    a = bf(inp)
    params = []
    for i, (w, b) in enumerate(ctr.cnn_split_params(p)):
        wt = torch.tensor(w, dtype=torch.float32, requires_grad=ste)
        bt = torch.tensor(b, dtype=torch.float32, requires_grad=ste)
        params += [wt, bt]
        z = F.conv3d(a, bf(wt), bt, padding=1)
        a = bf(torch.relu(z)) if i < 3 else z
    return a[0, 0], params


def _rel(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(37, 19, 11), (64, 16, 4), (8, 40, 5), (1, 7, 6), (17, 33, 34)])
def test_cnn_forward_matches_torch(ctx, shape):
    """Partial tiles in x and y, a one-voxel-wide volume (padded input rows
    of three voxels), and a z extent past one 32-slice run."""
    rng = np.random.default_rng(1)
    nx, ny, nz = shape
    p = ctr.cnn_init_params(3, 1.0)
    p[1728:1760] = rng.uniform(-0.1, 0.1, 32)            # nonzero biases
    p[-1] = 0.05
    x = rng.uniform(0, 1, nx * ny * nz).astype(np.float32)
    g = rng.standard_normal(nx * ny * nz).astype(np.float32)
    net = ctr.Cnn(ctx, p)
    out = net.forward(shape, x, g, 0.7)
    ref, _ = _torch_net(p, shape, x, g, 0.7)
    e = _rel(out, ref.detach().numpy())
    report(f"cnn_forward{shape}", rel_l2=e)
    assert e < FWD_TOL
    net.close()


@pytest.mark.gpu
def test_cnn_gradient_matches_torch(ctx):
    import torch
    rng = np.random.default_rng(2)
    shape, border, npatch = (36, 20, 9), 2, 2
    nx, ny, nz = shape
    n = nx * ny * nz
    p = ctr.cnn_init_params(4, 1.0)
    x = rng.uniform(0, 1, (npatch, n)).astype(np.float32)
    g = rng.standard_normal((npatch, n)).astype(np.float32)
    t = rng.standard_normal((npatch, n)).astype(np.float32) * 0.3
    gs = np.array([0.5, 1.5], np.float32)
    net = ctr.Cnn(ctx, p)
    loss, grad = net.train_step(shape, x, g, gs, t, border=border, want_grad=True)

    interior = (nx - 2 * border) * (ny - 2 * border) * (nz - 2 * border) * npatch
    total, grads = 0.0, None
    for i in range(npatch):
        out, params = _torch_net(p, shape, x[i], g[i], float(gs[i]), ste=True)
        tt = torch.from_numpy(t[i].reshape(nz, ny, nx))
        d = (out - tt)[border:nz - border, border:ny - border, border:nx - border]
        li = (d * d).sum() / interior
        li.backward()
        total += float(li)
        gi = np.concatenate([q.grad.numpy().ravel() for q in params])
        grads = gi if grads is None else grads + gi
    el = abs(loss - total) / total
    report("cnn_loss", rel=el)
    assert el < FWD_TOL
    o = 0
    for li, (co, ci) in enumerate(ctr.CNN_LAYERS):
        nw = co * ci * 27
        ew = _rel(grad[o:o + nw], grads[o:o + nw])
        eb = _rel(grad[o + nw:o + nw + co], grads[o + nw:o + nw + co])
        report(f"cnn_grad[layer {li + 1}]", w_rel_l2=ew, b_rel_l2=eb)
        assert ew < GRAD_TOL and eb < GRAD_TOL
        o += nw + co
    net.close()


@pytest.mark.gpu
def test_cnn_training_lowers_loss(ctx):
    rng = np.random.default_rng(5)
    shape = (24, 24, 12)
    n = 24 * 24 * 12
    x = rng.uniform(0, 1, (3, n)).astype(np.float32)
    g = rng.standard_normal((3, n)).astype(np.float32)
    t = (0.5 * g + 0.2 * x).astype(np.float32)
    gs = np.ones(3, np.float32)
    net = ctr.Cnn(ctx, ctr.cnn_init_params(6, 0.1))
    try:
        losses = [net.train_step(shape, x, g, gs, t, border=1, lr=3e-3) for _ in range(60)]
    finally:
        net.close()
    report("cnn_train", first=losses[0], last=losses[-1])
    assert losses[-1] < 0.6 * losses[0]
    assert all(np.isfinite(losses))


def _small_problem():
    g = geometry.cone_circular(32, 1.0, 24, 48, 48, 1.0, 100.0, 200.0)
    rng = np.random.default_rng(7)
    vol = np.zeros((32, 32, 32), np.float32)
    vol[8:24, 10:22, 6:26] = 1.0
    vol += rng.uniform(0, 0.2, vol.shape).astype(np.float32)
    return g, vol.ravel()


@pytest.mark.gpu
def test_learned_sirt_zero_network_is_plain_sirt(ctx):
    g, vol = _small_problem()
    b = ctx.cone_fp(g, vol)
    net = ctr.Cnn(ctx, ctr.cnn_init_params(0, 0.0))
    xl, rl, acc = ctx.learned_sirt3d(g, b, net, 6, lam=1.0)
    xp, rp, _ = ctx.sirt3d(g, b, 6, lam=1.0)
    e = _rel(xl, xp)
    report("learned_sirt[zero net]", x_rel=e, res_rel=_rel(rl, rp))
    assert e < 1e-5
    assert _rel(rl, rp) < 1e-5
    net.close()


@pytest.mark.gpu
def test_learned_sirt_residual_never_rises(ctx):
    g, vol = _small_problem()
    b = ctx.cone_fp(g, vol)
    net = ctr.Cnn(ctx, ctr.cnn_init_params(11, 1.0))   # an untrained, large correction
    x, r, acc = ctx.learned_sirt3d(g, b, net, 8, lam=1.0)
    report("learned_sirt[random net]", residuals=np.round(r, 5).tolist(), accepted=acc.tolist())
    assert np.all(np.diff(r) <= 1e-7)
    assert set(acc.tolist()) <= {0, 1}
    assert np.all(np.isfinite(x)) and x.min() >= 0.0
    net.close()


@pytest.mark.gpu
def test_learned_sirt_judged_against_plain_step(ctx):
    g, vol = _small_problem()
    b = ctx.cone_fp(g, vol)
    net = ctr.Cnn(ctx, ctr.cnn_init_params(12, 1.0))
    try:
        x, r, acc = ctx.learned_sirt3d(g, b, net, 6, lam=1.0, judge_vs_plain=True)
        _, rp, _ = ctx.sirt3d(g, b, 1, lam=1.0)
    finally:
        net.close()
    report("learned_sirt[judged vs plain]", residuals=np.round(r, 5).tolist(), accepted=acc.tolist())
    assert r[1] <= rp[1] * (1 + 1e-6)          # step 1 starts from the same iterate as the plain step
    assert np.all(np.diff(r) <= 1e-7)


@pytest.mark.gpu
@pytest.mark.parametrize("tile,halo", [(48, 8), (32, 4)])
def test_cnn_tiled_equals_whole_volume(ctx, tile, halo):
    """Tiles with halos of at least the receptive radius (4) and input boxes
    cropped at the volume faces give the whole-volume output bit for bit."""
    rng = np.random.default_rng(9)
    shape = (150, 90, 70)
    n = 150 * 90 * 70
    p = ctr.cnn_init_params(8, 1.0)
    p[1728:1760] = rng.uniform(-0.1, 0.1, 32)
    x = rng.uniform(0, 1, n).astype(np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    net = ctr.Cnn(ctx, p)
    try:
        whole = net.forward(shape, x, g, 0.8)
        tiled = net.forward_tiled(shape, x, g, 0.8, tile=tile, halo=halo)
    finally:
        net.close()
    report(f"cnn_tiled[{tile}, halo {halo}]", identical=bool(np.array_equal(whole, tiled)),
           max_abs=float(np.abs(whole - tiled).max()))
    assert np.array_equal(whole, tiled)


@pytest.mark.gpu
def test_learned_sirt_patchwise_equals_whole(ctx):
    g, vol = _small_problem()
    b = ctx.cone_fp(g, vol)
    net = ctr.Cnn(ctx, ctr.cnn_init_params(13, 0.5))
    try:
        x0, r0, a0 = ctx.learned_sirt3d(g, b, net, 4)
        x1, r1, a1 = ctx.learned_sirt3d(g, b, net, 4, cnn_tile=12)
    finally:
        net.close()
    report("learned_sirt[patch-wise]", identical=bool(np.array_equal(x0, x1)), accepted=a1.tolist())
    assert np.array_equal(x0, x1) and np.array_equal(r0, r1) and np.array_equal(a0, a1)


@pytest.mark.gpu
def test_emulated_sirt_is_exact_without_clamp(ctx):
    """Without the clamp and on the whole volume the emulation is linear; a
    Chebyshev degree of N - 1 interpolates the degree-(N-1) polynomial p_N
    exactly, so the result is plain unclamped SIRT of n_target iterations."""
    g, vol = _small_problem()
    b = ctx.cone_fp(g, vol)
    xe, res, fb = ctx.sirt3d_emulated(g, b, None, n_start=2, n_target=10, degree=7, n_polish=0, lo=-np.inf)
    xp, rp, _ = ctx.sirt3d(g, b, 10, lo=-np.inf)
    e = _rel(xe, xp)
    report("sirt3d_emulated[exact]", x_rel=e, residuals=np.round(res, 6).tolist(), plain=float(rp[-1]), fell_back=fb)
    assert e < 1e-4 and fb == 0
    assert abs(res[1] - rp[-1]) < 1e-4 * rp[-1]


@pytest.mark.gpu
def test_emulated_sirt_judged_by_residual(ctx):
    """A degree too low for the emulated count is still judged: the clamped
    result is kept only when it lowers the residual; otherwise the output is
    plain SIRT of n_start + n_polish iterations."""
    g, vol = _small_problem()
    b = ctx.cone_fp(g, vol)
    for degree, n_target in ((12, 60), (1, 400)):
        x, res, fb = ctx.sirt3d_emulated(g, b, None, n_start=3, n_target=n_target, degree=degree, n_polish=2)
        report(f"sirt3d_emulated[K={degree}, N={n_target}]", residuals=np.round(res, 6).tolist(), fell_back=fb)
        assert float(np.min(x)) >= 0.0
        if fb:
            xp, _, _ = ctx.sirt3d(g, b, 5)
            assert _rel(x, xp) < 1e-5
        else:
            assert res[1] < res[0] and res[2] <= res[1] * (1 + 1e-6)


@pytest.mark.gpu
def test_emulated_sirt_with_support_network_runs(ctx):
    g, vol = _small_problem()
    b = ctx.cone_fp(g, vol)
    net = ctr.Cnn(ctx, ctr.cnn_init_params(14, 0.5))
    try:
        x0, r0, f0 = ctx.sirt3d_emulated(g, b, net, n_start=3, n_target=40, degree=8, n_polish=1, threshold=0.0)
        x1, r1, f1 = ctx.sirt3d_emulated(g, b, net, n_start=3, n_target=40, degree=8, n_polish=1, threshold=0.0,
                                         cnn_tile=12)
    finally:
        net.close()
    report("sirt3d_emulated[network support]", residuals=np.round(r0, 6).tolist(), fell_back=f0,
           tiled_identical=bool(np.array_equal(x0, x1)))
    assert np.array_equal(x0, x1) and float(np.min(x0)) >= 0.0


@pytest.mark.gpu
def test_emulated_sirt_two_pass_support(ctx):
    """A first low-degree pass refines the support; the result stays judged and
    clamped, and an exact-degree two-pass run without the clamp is still plain
    unclamped SIRT (the first pass then leaves every voxel in the support)."""
    g, vol = _small_problem()
    b = ctx.cone_fp(g, vol)
    x, res, fb = ctx.sirt3d_emulated(g, b, None, n_start=3, n_target=60, degree=12, n_polish=1, support_degree=4)
    report("sirt3d_emulated[two-pass]", residuals=np.round(res, 6).tolist(), fell_back=fb)
    assert float(np.min(x)) >= 0.0
    assert fb == 1 or (res[1] < res[0] and res[2] <= res[1] * (1 + 1e-6))
    xe, _, _ = ctx.sirt3d_emulated(g, b, None, n_start=2, n_target=10, degree=7, n_polish=0, lo=-np.inf,
                                   support_degree=3)
    xp, _, _ = ctx.sirt3d(g, b, 10, lo=-np.inf)
    assert _rel(xe, xp) < 1e-4
