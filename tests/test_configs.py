"""The five benchmark workloads at their full sizes on the HIP path.

Each configuration runs the operators the reference's own callers chain
(FBP_fan_GPU.m:44-70: fan FP then FBP; FDK_GPU.m:74-106: cone FP then FDK;
SIRT3D.m:67-92: the plain loop) at the real problem size, and checks them
against the FP64 oracle where the oracle finishes in seconds (every view of
the 2-D case, view or slab subsets of the 3-D ones) and otherwise through
properties that do not depend on size:

  * FP of a constant-1 volume equals the chord of every ray through the
    volume box (slab method, FP64 on the device);
  * FDK of a centred uniform ball recovers its amplitude;
  * the adjoint identity <Ax, y> = <x, A'y> of the exact Siddon pair;
  * SIRT residuals never rise and iterates stay inside the clamp;
  * batched SIRT gives the iterates of one call per problem, bit for bit.

Sizes:
  C1  2-D fan 256^2 Shepp-Logan, 360 views, 512 equiangular bins (dgamma 1e-3, SOD 500)
  C2  256^3 at 1 mm, 360 views, 512^2 flat detector at 0.5 mm iso pitch, SOD 600
  C3  16 seeded problems N in {128..320}, views in {90,180,360}, detectors 384..640
  C4  256^3, 90 views, Poisson noise at I0 = 1e5, the learned step
  C5  512^3 at 0.5 mm, 720 views, 1024 x 768 detector at 0.509 mm iso pitch, SOD 700
"""
import time

import numpy as np
import pytest

import ctr
from ctr import geometry, learned, phantoms
from conftest import errs, report

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def dev(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def cone_chords(g, views, device="cuda"):
    """Chord (mm) of every ray (beta, xi, p) of the given views through the
    volume box [0, size], from the source S(beta) through the isocentre
    detector point, by the slab method in FP64 on the device."""
    torch = _torch()
    f64 = dict(dtype=torch.float64, device=device)
    b = torch.tensor(g.beta[views].astype(np.float64), **f64)[:, None, None]
    p = torch.tensor(g.p.astype(np.float64), **f64)[None, None, :]
    xi = torch.tensor(g.xi.astype(np.float64), **f64)[None, :, None]
    D = g.distance
    sb, cb = torch.sin(b), torch.cos(b)
    S = [g.size[0] / 2 - D * sb, g.size[1] / 2 + D * cb, torch.full_like(b, g.size[2] / 2)]
    d = [p * cb + D * sb, p * sb - D * cb, xi + 0 * b]
    shape = torch.broadcast_shapes(*(x.shape for x in d))
    lo = torch.full(shape, -np.inf, **f64)
    hi = torch.full(shape, np.inf, **f64)
    norm = torch.sqrt(sum(x * x for x in d)).expand(shape)
    for a in range(3):
        da = d[a].expand(shape)
        sa = S[a].expand(shape)
        nz = da != 0
        t1 = torch.where(nz, (0.0 - sa) / torch.where(nz, da, 1.0), -np.inf)
        t2 = torch.where(nz, (g.size[a] - sa) / torch.where(nz, da, 1.0), np.inf)
        lo = torch.maximum(lo, torch.minimum(t1, t2))
        hi = torch.minimum(hi, torch.maximum(t1, t2))
        inside = (sa >= 0) & (sa <= g.size[a])
        lo = torch.where(nz | inside, lo, np.inf)        # parallel to the slab and outside it: no chord
    return (torch.clamp(hi - lo, min=0) * norm).reshape(-1)


def adjoint_gap(ctx, g, seed):
    """|<Ax, y> - <x, A'y>| / (|Ax| |y|) of the exact pair, FP64 dot products."""
    torch = _torch()
    gen = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.rand(g.nvox, device="cuda", generator=gen)
    y = torch.rand(g.nray, device="cuda", generator=gen)
    ax = ctx.cone_fp(g, x)
    aty = ctx.cone_bp_adjoint(g, y)
    lhs = torch.dot(ax.double(), y.double())
    rhs = torch.dot(x.double(), aty.double())
    return float(abs(lhs - rhs) / (torch.linalg.norm(ax.double()) * torch.linalg.norm(y.double())))


def ball(n, voxel, radius_mm, amp=1.0):
    torch = _torch()
    c = (torch.arange(n, device="cuda", dtype=torch.float64) + 0.5 - n / 2) * voxel
    r2 = c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2
    return (amp * (r2 <= radius_mm ** 2)).to(torch.float32).reshape(-1), r2.sqrt()


def edge_columns(g, views=slice(None), tol_mm=1e-3):
    """[ns, nt] mask of the voxel columns whose centre projects within tol_mm
    of the first or last detector column in some view (FP64).  On whether
    such a sample is on the detector, the HIP kernels (float) and the FP64
    oracle can decide differently, which adds or drops one edge sample."""
    torch = _torch()
    f64 = dict(dtype=torch.float64, device="cuda")
    rt, rs = g.size[0] / g.nt, g.size[1] / g.ns
    x = ((torch.arange(g.nt, **f64) + 0.5) * rt - g.size[0] / 2)[None, :]
    y = ((torch.arange(g.ns, **f64) + 0.5) * rs - g.size[1] / 2)[:, None]
    ends = (float(g.p[0]), float(g.p[-1]))
    near = torch.zeros((g.ns, g.nt), dtype=torch.bool, device="cuda")
    for b in g.beta[views].astype(np.float64):
        u = x * np.cos(b) + y * np.sin(b)
        w = -x * np.sin(b) + y * np.cos(b)
        P = u * g.distance / (g.distance - w)
        near |= ((P - ends[0]).abs() < tol_mm) | ((P - ends[1]).abs() < tol_mm)
    return near


def masked_errs(x, ref, mask):
    """errs() over all voxels, and the max-relative error outside mask (a
    [ns, nt] column mask broadcast over the slices)."""
    e, m = errs(x, ref)
    x, ref = np.asarray(x, np.float64), np.asarray(ref, np.float64)
    keep = ~np.broadcast_to(mask, x.shape)
    m_in = float(np.abs(x - ref)[keep].max() / np.abs(ref).max())
    return e, m, m_in


def err_summary(x, ref):
    """Where the largest deviation sits and how the deviations are spread."""
    x, ref = np.asarray(x, np.float64).ravel(), np.asarray(ref, np.float64).ravel()
    d = np.abs(x - ref) / np.abs(ref).max()
    k = int(np.argmax(d))
    return dict(argmax=k, x=float(x[k]), ref=float(ref[k]), p99=float(np.quantile(d, 0.99)),
                p9999=float(np.quantile(d, 0.9999)))


def sub_cone(g, views):
    """The same geometry restricted to a subset of its views."""
    return ctr.ConeGeom(g.nt, g.ns, g.nz, g.size, g.distance, g.beta[views], g.p, g.xi)


def oracle_cone(O, g, views=slice(None), nz=None):
    """Oracle geometry; nz < g.nz keeps the central nz slices (same voxel
    centres: the volume is centred on the rotation axis)."""
    nz = g.nz if nz is None else nz
    rz = g.size[2] / g.nz
    return O.Cone(g.nt, g.ns, nz, (g.size[0], g.size[1], rz * nz), g.distance, g.beta[views], g.p, g.xi)


# ----------------------------------------------------------------------------
# C1: 2-D fan beam, 256^2, 360 views, 512 equiangular bins
# ----------------------------------------------------------------------------
class TestC1Fan:
    @pytest.fixture(scope="class")
    def c1(self, oracle):
        g = geometry.config_fan_2d()
        og = oracle.Fan(g.nt, g.ns, g.size, g.distance, g.beta, g.p)
        img = phantoms.shepp_logan_2d(256).astype(np.float32).ravel()
        print(f"[case] C1 fan {g.nt}x{g.ns}, {len(g.beta)} views, {len(g.p)} bins", flush=True)
        return g, og, img

    def test_fp_matches_oracle(self, ctx, oracle, c1):
        g, og, img = c1
        gpu = ctx.fan_fp(g, img)
        e1, m1 = errs(gpu, oracle.fan_fp_exact(og, img))
        e2, m2 = errs(gpu, oracle.fan_fp_ref(og, img))
        report("C1 fan_fp", rel_l2_exact=e1, max_rel_exact=m1, rel_l2_ref=e2, max_rel_ref=m2)
        assert e1 < 2e-6 and m1 < 2e-5
        assert e2 < 2e-6 and m2 < 2e-5

    @pytest.mark.parametrize("filt", [ctr.FILTER_RL_HALF, ctr.FILTER_SL_HALF])
    def test_equiangular_fbp_matches_oracle(self, ctx, oracle, c1, filt):
        g, og, img = c1
        proj = ctx.fan_fp(g, img)
        f_gpu = ctx.fan_filter(g, proj, ctr.FAN_EQUIANGULAR, filt)
        f_or = oracle.fan_filter(og, proj, True, filt)
        ef, mf = errs(f_gpu, f_or)
        rec = ctx.fan_fbp(g, proj, ctr.FAN_EQUIANGULAR, filt)
        with oracle.interp_mode(1):
            ref = oracle.fan_fbp_bp(og, f_or.astype(np.float32), True)
        e, m = errs(rec, ref)
        q = np.linalg.norm(rec - img) / np.linalg.norm(img)
        report(f"C1 fan_fbp[equiangular, filter {filt}]", filter_rel_l2=ef, filter_max_rel=mf, rel_l2=e, max_rel=m,
               vs_phantom=float(q))
        assert ef < 1e-6 and mf < 1e-5
        assert e < 2e-5 and m < 2e-4
        assert q < 0.3              # sampled Shepp-Logan: edges dominate the error

    def test_adjoint_and_sirt2d(self, ctx, oracle, c1):
        g, og, img = c1
        rng = np.random.default_rng(11)
        x = rng.random(g.nvox).astype(np.float32)
        y = rng.random(g.nray).astype(np.float32)
        ax = ctx.fan_fp(g, x).astype(np.float64)
        aty = ctx.fan_bp_adjoint(g, y).astype(np.float64)
        gap = abs(ax @ y - x @ aty) / (np.linalg.norm(ax) * np.linalg.norm(y))
        ea, ma = errs(aty, oracle.fan_adjoint(og, y))
        b = ctx.fan_fp(g, img)
        xg, r2g, _ = ctx.sirt2d(g, b, 5, lam=1.0, lo=0.0, hi=1.0)
        xo, r2o, _ = oracle.sirt2d(og, b, 5, lam=1.0, lo=0.0, hi=1.0)
        es, ms = errs(xg, xo)
        report("C1 adjoint/sirt2d", adjoint_gap=float(gap), adjoint_rel_l2=ea, sirt_rel_l2=es, sirt_max_rel=ms,
               res_gpu=float(r2g[-1]), res_oracle=float(r2o[-1]))
        assert gap < 1e-5
        assert ea < 1e-5 and ma < 1e-4
        assert es < 2e-5 and ms < 2e-4
        assert np.all(np.diff(r2g) <= 1e-7)


# ----------------------------------------------------------------------------
# C2: 256^3, 360 views, 512^2 detector
# ----------------------------------------------------------------------------
class TestC2Cone:
    @pytest.fixture(scope="class")
    def c2(self):
        g = geometry.config_cone_256(360)
        vol = dev(phantoms.phantom3d(256).ravel())
        print(f"[case] C2 cone {g.vol_shape}, proj {g.proj_shape}", flush=True)
        return g, vol

    def test_fp_matches_oracle_on_view_subset(self, ctx, oracle, c2):
        g, vol = c2
        proj = ctx.cone_fp(g, vol).view(g.proj_shape)
        sel = slice(0, 360, 45)
        og = oracle_cone(oracle, g, sel)
        v = host(vol)
        e1, m1 = errs(host(proj[sel]), oracle.cone_fp_exact(og, v))
        og1 = oracle_cone(oracle, g, slice(7, 8))
        e2, m2 = errs(host(proj[7:8]), oracle.cone_fp_ref(og1, v))
        report("C2 cone_fp[8 of 360 views]", rel_l2_exact=e1, max_rel_exact=m1, rel_l2_ref_view7=e2, max_rel_ref=m2)
        assert e1 < 2e-6 and m1 < 2e-5
        assert e2 < 2e-6 and m2 < 2e-5

    def test_fp_of_ones_is_box_chord(self, ctx, c2):
        torch = _torch()
        g, _ = c2
        ones = torch.ones(g.nvox, device="cuda")
        fp = ctx.cone_fp(g, ones).view(len(g.beta), -1)
        worst, num, den = 0.0, 0.0, 0.0
        for v0 in range(0, len(g.beta), 60):
            views = slice(v0, v0 + 60)
            ch = cone_chords(g, views)
            d = fp[views].reshape(-1).double() - ch
            worst = max(worst, float(d.abs().max()))
            num += float((d * d).sum())
            den += float((ch * ch).sum())
        rel = (num / den) ** 0.5
        report("C2 fp(ones) vs chord", max_abs_mm=worst, rel_l2=rel)
        assert rel < 1e-6 and worst < 2e-3        # chords up to ~440 mm

    def test_filter_matches_oracle_on_view_subset(self, ctx, oracle, c2):
        """The FFT ramp filter (pre-weight, SL-half and RL-half kernels) at
        full size against the oracle's FP64 direct convolution on 8 views."""
        g, vol = c2
        proj = ctx.cone_fp(g, vol).view(g.proj_shape)
        sel = slice(3, 360, 45)
        og = oracle_cone(oracle, g, sel)
        ph = host(proj[sel])
        for filt in (ctr.FILTER_SL_HALF, ctr.FILTER_RL_HALF):
            gpu = host(ctx.cone_filter(g, proj, filt=filt).view(g.proj_shape)[sel])
            e, m = errs(gpu, oracle.cone_filter(og, ph, filt=filt))
            report(f"C2 cone_filter[{filt}, 8 of 360 views]", rel_l2=e, max_rel=m)
            assert e < 1e-6 and m < 1e-5

    def test_fdk_bp_matches_oracle(self, ctx, oracle, c2):
        """The Shepp-Logan shadow is truncated by the detector in u at some
        views, so the filtered samples at the detector's u edges are large;
        columns that project within 1e-3 mm of an edge are kept out of the
        max-error bound (counted and reported) and stay in the rel-L2."""
        torch = _torch()
        g, vol = c2
        proj = ctx.cone_fp(g, vol)
        filt = ctx.cone_filter(g, proj, ctr.FILTER_SL_HALF)
        rec = ctx.fdk_bp(g, filt).view(g.vol_shape)
        fh = host(filt)
        # all 360 views on the central 16 slices
        og = oracle_cone(oracle, g, nz=16)
        with oracle.interp_mode(1):
            ref = oracle.fdk_bp(og, fh).reshape(16, 256, 256)
        edge = host(edge_columns(g))
        e1, m1, m1_in = masked_errs(host(rec[120:136]), ref, edge)
        # the full volume (every footprint at the edges of the detector) on 45 views
        sel = slice(0, 360, 8)
        gs = sub_cone(g, sel)
        fs = filt.view(g.proj_shape)[sel].contiguous()
        rec45 = ctx.fdk_bp(gs, fs)
        with oracle.interp_mode(1):
            ref45 = oracle.fdk_bp(oracle_cone(oracle, g, sel), host(fs)).reshape(g.vol_shape)
        edge45 = host(edge_columns(g, sel))
        e2, m2, m2_in = masked_errs(host(rec45).reshape(g.vol_shape), ref45, edge45)
        report("C2 fdk_bp", rel_l2_slab_360=e1, max_rel_slab_360=m1, max_rel_off_edge_360=m1_in,
               edge_columns_360=int(edge.sum()), rel_l2_full_45=e2, max_rel_full_45=m2, max_rel_off_edge_45=m2_in,
               edge_columns_45=int(edge45.sum()), **{f"slab_{k}": v for k, v in err_summary(host(rec[120:136]), ref).items()})
        assert e1 < 3e-5 and m1_in < 1e-4
        assert e2 < 3e-5 and m2_in < 1e-4
        assert edge.sum() < 0.01 * edge.size
        del rec, filt, proj
        torch.cuda.empty_cache()

    def test_fdk_of_ball_recovers_amplitude(self, ctx, c2):
        g, _ = c2
        b, r = ball(256, 1.0, 100.0)
        rec = ctx.fdk(g, ctx.cone_fp(g, b), ctr.FILTER_RL_HALF).view(256, 256, 256).double()
        z = r.new_tensor((np.arange(256) + 0.5 - 128)).view(256, 1, 1).expand_as(r)
        core = float(rec[(r < 60) & (z.abs() < 30)].mean())
        out = float(rec[(r > 110) & (r < 125)].mean())
        report("C2 fdk(ball r=100mm)", core_mean=core, shell_mean=out)
        assert abs(core - 1.0) < 0.01
        assert abs(out) < 0.01

    def test_bp_ref_and_adjoint_match_oracle(self, ctx, oracle, c2):
        """The reference BP and the exact adjoint on 8 views against the
        oracle, and the adjoint identity over all 360.  The adjoint's
        interval ends are float parameters measured from the source, 600 mm
        away (about 6e-5 mm each), which bounds its agreement with the FP64
        oracle here at rel-L2 2e-5 and 1e-4 of the largest voxel."""
        g, _ = c2
        sel = slice(3, 360, 45)
        gs = sub_cone(g, sel)
        og = oracle_cone(oracle, g, sel)
        y = np.random.default_rng(21).random(gs.nray).astype(np.float32)
        yd = dev(y)
        with oracle.interp_mode(1):
            e1, m1 = errs(host(ctx.cone_bp_ref(gs, yd)), oracle.cone_bp_ref(og, y))
        adj = host(ctx.cone_bp_adjoint(gs, yd))
        adj_ref = oracle.cone_adjoint(og, y)
        e2, m2 = errs(adj, adj_ref)
        gap = adjoint_gap(ctx, g, 5)
        report("C2 bp", bp_ref_rel_l2=e1, bp_ref_max_rel=m1, adjoint_rel_l2=e2, adjoint_max_rel=m2,
               adjoint_gap_360=gap, **{f"adj_{k}": v for k, v in err_summary(adj, adj_ref).items()})
        assert e1 < 1e-5 and m1 < 1e-4
        assert e2 < 2e-5 and m2 < 1e-4
        assert gap < 1e-5

    def test_sirt_three_iterations(self, ctx, c2):
        torch = _torch()
        g, vol = c2
        b = ctx.cone_fp(g, vol)
        x, r2, r1 = ctx.sirt3d(g, b, 3)
        # the first iterate from zero is clamp(V A'(W b)), assembled from the operators
        w = 1.0 / (ctx.cone_fp(g, torch.ones(g.nvox, device="cuda")) + 1e-9)
        v = 1.0 / (ctx.cone_bp_adjoint(g, torch.ones(g.nray, device="cuda")) + 1e-9)
        x1_ops = torch.clamp(v * ctx.cone_bp_adjoint(g, w * b), min=0)
        x1, _, _ = ctx.sirt3d(g, b, 1)
        e1 = float(torch.linalg.norm(x1 - x1_ops) / torch.linalg.norm(x1_ops))
        report("C2 sirt3d[3 iters]", residuals=np.round(r2, 6).tolist(), l1=np.round(r1, 9).tolist(),
               first_iterate_vs_operators=e1, x_min=float(x.min()))
        assert np.all(np.diff(r2) < 0)
        assert float(x.min()) >= 0.0 and bool(torch.isfinite(x).all())
        assert e1 < 1e-5


# ----------------------------------------------------------------------------
# C3: 16 uneven problems, batched
# ----------------------------------------------------------------------------
def test_c3_batched_matches_single(ctx):
    torch = _torch()
    from ctr import batched
    probs = batched.make_problems(ctx, 16)
    print("[case] C3 " + ", ".join(f"{n}^3/{v}v/{d}px" for _, _, n, v, d in probs), flush=True)
    n_iter = 2
    xs, res = ctx.sirt3d_batched([(g, b) for g, b, *_ in probs], n_iter)
    worst, worst_r = 0.0, 0.0
    for i, (g, b, n, v, d) in enumerate(probs):
        x1, r1, _ = ctx.sirt3d(g, b, n_iter)
        same = bool(torch.equal(xs[i], x1))
        dr = float(np.abs(res[i] - r1).max())
        worst = max(worst, float((xs[i] - x1).abs().max()))
        worst_r = max(worst_r, dr)
        assert same, f"problem {i} ({n}^3, {v} views, {d}px) differs from its single run"
        assert np.all(np.diff(res[i]) < 0)
    report("C3 batched vs single", problems=len(probs), max_dx=worst, max_dres=worst_r)
    assert worst_r <= 1e-9
    del xs, probs
    torch.cuda.empty_cache()


# ----------------------------------------------------------------------------
# C4: 256^3, 90 views, Poisson noise; the learned step
# ----------------------------------------------------------------------------
class TestC4Learned:
    @pytest.fixture(scope="class")
    def c4(self, ctx):
        g, vol, b = learned.make_problem(ctx, 2001)
        print(f"[case] C4 {g.vol_shape}, proj {g.proj_shape}, Poisson I0=1e5", flush=True)
        return g, vol, b

    def test_zero_network_is_plain_sirt(self, ctx, c4):
        torch = _torch()
        g, _, b = c4
        net = ctr.Cnn(ctx, ctr.cnn_init_params(0, 0.0))
        try:
            xl, rl, acc = ctx.learned_sirt3d(g, b, net, 3)
        finally:
            net.close()
        xp, rp, _ = ctx.sirt3d(g, b, 3)
        e = float(torch.linalg.norm(xl - xp) / torch.linalg.norm(xp))
        report("C4 learned[zero net]", x_rel=e, res_l=np.round(rl, 6).tolist(), res_p=np.round(rp, 6).tolist())
        assert e < 1e-5
        assert np.allclose(rl, rp, rtol=1e-5, atol=0)

    def test_trained_network_never_raises_residual(self, ctx, c4):
        torch = _torch()
        g, _, b = c4
        with np.load(learned.WEIGHTS, allow_pickle=False) as f:
            params = f["params"]
        net = ctr.Cnn(ctx, params)
        try:
            x, r, acc = ctx.learned_sirt3d(g, b, net, 4, judge_vs_plain=True)
        finally:
            net.close()
        _, rp, _ = ctx.sirt3d(g, b, 1)
        report("C4 learned[trained]", residuals=np.round(r, 6).tolist(), accepted=acc.tolist())
        assert np.all(np.diff(r) <= 1e-7)
        assert r[1] <= rp[1] * (1 + 1e-6)
        assert float(x.min()) >= 0.0 and bool(torch.isfinite(x).all())

    def test_cnn_full_volume_matches_torch_on_subblock(self, ctx, c4):
        """The network over the whole 256^3 volume, checked on a 40^3 block
        against a PyTorch fp32 reference that sees the block plus the
        network's 4-voxel receptive-field halo."""
        import test_cnn
        torch = _torch()
        g, vol, b = c4
        x, _, _ = ctx.sirt3d(g, b, 5)
        s = ctx.sirt3d_step(g, b, x)
        rms = float(torch.sqrt(torch.mean(s * s)))
        with np.load(learned.WEIGHTS, allow_pickle=False) as f:
            params = f["params"]
        net = ctr.Cnn(ctx, params)
        try:
            out = net.forward((256, 256, 256), x, s, 1.0 / rms).view(256, 256, 256)
        finally:
            net.close()
        z0, y0, x0, m, h = 100, 30, 180, 40, 4
        blk = (slice(z0 - h, z0 + m + h), slice(y0 - h, y0 + m + h), slice(x0 - h, x0 + m + h))
        xb = host(x.view(256, 256, 256)[blk]).ravel()
        sb = host(s.view(256, 256, 256)[blk]).ravel()
        ref, _ = test_cnn._torch_net(params, (m + 2 * h,) * 3, xb, sb, 1.0 / rms)
        ref = ref.detach().numpy()[h:-h, h:-h, h:-h]
        got = host(out[z0:z0 + m, y0:y0 + m, x0:x0 + m])
        e = test_cnn._rel(got, ref)
        report("C4 cnn[256^3, 40^3 block]", rel_l2=e)
        assert e < test_cnn.FWD_TOL


# ----------------------------------------------------------------------------
# C5: 512^3, 720 views, 1024 x 768 detector
# ----------------------------------------------------------------------------
class TestC5Large:
    @pytest.fixture(scope="class")
    def c5(self):
        torch = _torch()
        g = geometry.config_cone_512()
        vol = phantoms.rasterize_3d_torch(phantoms.random_ellipsoids(5005), 512).reshape(-1).contiguous()
        print(f"[case] C5 cone {g.vol_shape}, proj {g.proj_shape}", flush=True)
        yield g, vol
        del vol
        torch.cuda.empty_cache()

    def test_fp_matches_oracle_and_chord(self, ctx, oracle, c5):
        """The volume is above 256 MiB, so the projector takes its masked-read
        path; two views against the oracle, every ray of FP(1) against the chord."""
        torch = _torch()
        g, vol = c5
        t0 = time.perf_counter()
        proj = ctx.cone_fp(g, vol).view(g.proj_shape)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0)
        sel = slice(100, 720, 360)
        e1, m1 = errs(host(proj[sel]), oracle.cone_fp_exact(oracle_cone(oracle, g, sel), host(vol)))
        del proj
        fp1 = ctx.cone_fp(g, torch.ones(g.nvox, device="cuda")).view(len(g.beta), -1)
        worst, num, den = 0.0, 0.0, 0.0
        for v0 in range(0, 720, 40):
            views = slice(v0, v0 + 40)
            ch = cone_chords(g, views)
            d = fp1[views].reshape(-1).double() - ch
            worst = max(worst, float(d.abs().max()))
            num += float((d * d).sum())
            den += float((ch * ch).sum())
        rel = (num / den) ** 0.5
        report("C5 cone_fp", first_call_ms=ms, rel_l2_2views=e1, max_rel_2views=m1, ones_rel_l2=rel,
               ones_max_abs_mm=worst)
        assert e1 < 2e-6 and m1 < 2e-5
        assert rel < 1e-6 and worst < 2e-3
        del fp1
        torch.cuda.empty_cache()

    def test_fdk_ball_and_slab(self, ctx, oracle, c5):
        torch = _torch()
        g, _ = c5
        b, r = ball(512, 0.5, 100.0)
        proj = ctx.cone_fp(g, b)
        filt = ctx.cone_filter(g, proj, ctr.FILTER_RL_HALF)
        del proj
        rec = ctx.fdk_bp(g, filt).view(512, 512, 512)
        z = r.new_tensor((np.arange(512) + 0.5 - 256) * 0.5).view(512, 1, 1).expand_as(r)
        core = float(rec.double()[(r < 60) & (z.abs() < 30)].mean())
        shell = float(rec.double()[(r > 110) & (r < 125)].mean())
        og = oracle_cone(oracle, g, nz=4)
        with oracle.interp_mode(1):
            ref = oracle.fdk_bp(og, host(filt)).reshape(4, 512, 512)
        e, m = errs(host(rec[254:258]), ref)
        report("C5 fdk", core_mean=core, shell_mean=shell, rel_l2_slab=e, max_rel_slab=m)
        assert abs(core - 1.0) < 0.01 and abs(shell) < 0.01
        assert e < 3e-5 and m < 1e-4
        del rec, filt, b, r
        torch.cuda.empty_cache()

    def test_adjoint_identity(self, ctx, c5):
        g, _ = c5
        gap = adjoint_gap(ctx, g, 9)
        report("C5 adjoint", gap=gap)
        assert gap < 1e-5

    def test_patchwise_cnn_and_learned_steps(self, ctx, c5):
        """The network over the 512^3 volume in 128^3 tiles with 8-voxel halos
        equals the whole-volume pass bit for bit, and the learned loop runs
        patch-wise at this size with the residual never rising."""
        torch = _torch()
        g, vol = c5
        b = ctx.cone_fp(g, vol)
        x, _, _ = ctx.sirt3d(g, b, 2)
        s = ctx.sirt3d_step(g, b, x)
        rms = float(torch.sqrt(torch.mean(s * s)))
        with np.load(learned.WEIGHTS, allow_pickle=False) as f:
            params = f["params"]
        net = ctr.Cnn(ctx, params)
        try:
            tiled = net.forward_tiled((512, 512, 512), x, s, 1.0 / rms, tile=128, halo=8)
            whole = net.forward((512, 512, 512), x, s, 1.0 / rms)
            same = bool(torch.equal(tiled, whole))
            del whole
            torch.cuda.empty_cache()
            t0 = time.perf_counter()
            xl, rl, acc = ctx.learned_sirt3d(g, b, net, 2, judge_vs_plain=True, cnn_tile=128)
            torch.cuda.synchronize()
            ms = 1e3 * (time.perf_counter() - t0)
        finally:
            net.close()
        report("C5 patch-wise cnn", identical=same, learned_residuals=np.round(rl, 6).tolist(),
               accepted=acc.tolist(), learned_2_steps_ms=ms)
        assert same
        assert np.all(np.diff(rl) <= 1e-7) and float(xl.min()) >= 0.0
        del xl, x, s, b, tiled
        torch.cuda.empty_cache()
