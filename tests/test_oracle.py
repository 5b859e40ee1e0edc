"""Pins the FP64 oracle (oracle/ctr_oracle.c) before it is trusted as the
checker: brute force on tiny inputs, closed forms, invariants and an
independent dense-matrix SIRT.  CPU only."""
import numpy as np
import pytest

from analytic import curved_ball_projections


def small_cone(O, nt=7, ns=6, nz=5, size=(7.0, 6.6, 4.5), D=25.0, nb=9, np_=11, nx=9, pp=0.9, px=0.7,
               beta0=None):
    beta = (2 * np.pi / nb) * np.arange(1, nb + 1) if beta0 is None else np.asarray(beta0)
    p = pp * (np.arange(np_) - (np_ - 1) / 2)
    xi = px * (np.arange(nx) - (nx - 1) / 2)
    return O.Cone(nt, ns, nz, size, D, beta.astype(np.float32), p.astype(np.float32), xi.astype(np.float32))


def rays(og):
    """Source and unit direction of every ray, computed independently in numpy."""
    beta = og.beta.astype(np.float64)[:, None, None]
    p = og.p.astype(np.float64)[None, None, :]
    xi = og.xi.astype(np.float64)[None, :, None]
    D = og.s.D
    size = np.array(og.s.size[:])
    S = np.stack(np.broadcast_arrays(size[0] / 2 - D * np.sin(beta), size[1] / 2 + D * np.cos(beta),
                                     size[2] / 2 + 0 * beta), -1)
    d = np.stack(np.broadcast_arrays(p * np.cos(beta) + D * np.sin(beta), p * np.sin(beta) - D * np.cos(beta),
                                     xi + 0 * beta), -1)
    d = d / np.linalg.norm(d, axis=-1, keepdims=True)
    S = np.broadcast_to(S, d.shape)
    return S.reshape(-1, 3), d.reshape(-1, 3), size


def test_siddon_exact_matches_brute_force(oracle):
    rng = np.random.default_rng(0)
    for kw in [{}, dict(size=(7.0, 7.0, 7.0), nz=7), dict(D=60.0, pp=0.3, px=0.25, nb=5)]:
        og = small_cone(oracle, **kw)
        vol = rng.random(og.nvox).astype(np.float32)
        ex = oracle.cone_fp_exact(og, vol)
        nb, nx, np_ = len(og.beta), len(og.xi), len(og.p)
        br = np.array([oracle.cone_ray_brute(og, vol, b, ip, ix) for b in range(nb) for ix in range(nx)
                       for ip in range(np_)])
        assert np.abs(ex - br).max() <= 1e-10 * np.abs(br).max()


def test_fp_of_ones_is_box_chord(oracle):
    """FP of a constant-1 volume = chord of the ray through the volume box (slab method)."""
    og = small_cone(oracle, nt=9, ns=8, nz=6, size=(9.0, 8.8, 5.4), np_=15, nx=13, pp=1.1, px=0.9)
    S, d, size = rays(og)
    with np.errstate(divide="ignore", invalid="ignore"):
        t1 = (0 - S) / d
        t2 = (size - S) / d
    lo = np.nanmax(np.where(d != 0, np.minimum(t1, t2), -np.inf), axis=1)
    hi = np.nanmin(np.where(d != 0, np.maximum(t1, t2), np.inf), axis=1)
    chord = np.maximum(hi - lo, 0)
    fp1 = oracle.cone_fp_exact(og, np.ones(og.nvox, np.float32))
    assert np.abs(fp1 - chord).max() < 1e-10
    assert np.abs(oracle.cone_rowsum(og) - chord).max() < 1e-10
    assert (chord > 0).sum() > 0.25 * len(chord)


def test_fp_of_ball_converges_to_analytic_chord(oracle):
    """Line integrals of a rasterised ball approach 2 sqrt(r^2 - dist^2) (ray-sphere closed form)."""
    n = 96
    size = (48.0, 48.0, 48.0)
    og = small_cone(oracle, nt=n, ns=n, nz=n, size=size, D=200.0, nb=3, np_=21, nx=5, pp=1.7, px=3.1)
    c = np.array(size) / 2
    r = 15.0
    g = (np.arange(n) + 0.5) * 0.5
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")   # [z][s][t]
    vol = (((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) <= r * r).astype(np.float32).ravel()
    fp = oracle.cone_fp_exact(og, vol)
    S, d, _ = rays(og)
    w = S - c
    tca = -(w * d).sum(1)
    dist2 = (w * w).sum(1) - tca ** 2
    chord = 2 * np.sqrt(np.maximum(r * r - dist2, 0))
    inside = chord > 10
    assert inside.sum() > 50
    # staircase boundary of a 0.5 mm raster: mean error well under a voxel, totals to 0.3 %
    assert np.abs(fp[inside] - chord[inside]).mean() < 0.25
    assert abs(fp[inside].sum() / chord[inside].sum() - 1) < 3e-3


def test_ref_restatement_agrees_with_exact(oracle):
    rng = np.random.default_rng(1)
    og = small_cone(oracle, nt=12, ns=11, nz=7, size=(12.0, 12.1, 6.3), np_=23, nx=15, pp=0.8, px=0.6)
    vol = rng.random(og.nvox).astype(np.float32)
    ex, rf = oracle.cone_fp_exact(og, vol), oracle.cone_fp_ref(og, vol)
    assert np.abs(ex - rf).max() < 1e-7 * np.abs(ex).max()
    # no ray lies exactly in a voxel plane (those are assigned to either side by rounding)
    fg = oracle.Fan(12, 11, (12.0, 12.1), 30.0, (2 * np.pi / 13) * np.arange(1, 14), 0.8 * (np.arange(25) - 12) + 0.1)
    img = rng.random(fg.nvox).astype(np.float32)
    assert np.abs(oracle.fan_fp_exact(fg, img) - oracle.fan_fp_ref(fg, img)).max() < 1e-7 * 12


def test_mirror_symmetry(oracle):
    """At beta = 0 the t-mirror of the volume maps ray p to ray -p, the z-mirror maps xi to -xi."""
    # even sample counts: no ray lies in the mirror plane itself
    og = small_cone(oracle, nt=10, ns=9, nz=8, size=(10.0, 9.0, 8.0), beta0=[0.0], np_=16, nx=12, pp=0.9, px=0.8)
    vol = np.random.default_rng(2).random(og.nvox).astype(np.float32).reshape(8, 9, 10)
    a = oracle.cone_fp_exact(og, vol.ravel()).reshape(12, 16)
    b = oracle.cone_fp_exact(og, vol[:, :, ::-1].ravel()).reshape(12, 16)
    c = oracle.cone_fp_exact(og, vol[::-1].ravel()).reshape(12, 16)
    assert np.allclose(a, b[:, ::-1], rtol=1e-12, atol=1e-12)
    assert np.allclose(a, c[::-1, :], rtol=1e-12, atol=1e-12)


def test_adjoint_identity(oracle):
    rng = np.random.default_rng(3)
    og = small_cone(oracle)
    x = rng.random(og.nvox).astype(np.float32)
    y = rng.random(og.nray).astype(np.float32)
    lhs = np.dot(oracle.cone_fp_exact(og, x), y.astype(np.float64))
    rhs = np.dot(x.astype(np.float64), oracle.cone_adjoint(og, y))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    fg = oracle.Fan(9, 8, (9.0, 8.8), 30.0, (2 * np.pi / 11) * np.arange(1, 12), 0.7 * (np.arange(19) - 9))
    x2 = rng.random(fg.nvox).astype(np.float32)
    y2 = rng.random(fg.nray).astype(np.float32)
    assert abs(np.dot(oracle.fan_fp_exact(fg, x2), y2) - np.dot(x2, oracle.fan_adjoint(fg, y2))) < 1e-10


def _response(h, w):
    n = np.arange(1, len(h))
    return h[0] + 2 * np.sum(h[1:, None] * np.cos(np.outer(n, w)), axis=0)


def test_filter_taps_frequency_response(oracle):
    """Half Shepp-Logan: H(w) = |sin(w/2)|/(2 pi); half Ram-Lak: H(w) = |w|/(4 pi), |w| <= pi:
    both are half the ramp |w|/(2 pi) at low frequency (d = 1 here; the taps scale as
    1/d^2 and the convolution multiplies by d)."""
    N = 200000
    w = np.linspace(0.05, np.pi, 9)
    sl = oracle.filter_taps(0, False, 1.0, N)
    rl = oracle.filter_taps(1, False, 1.0, N)
    assert np.allclose(_response(sl, w), np.abs(np.sin(w / 2)) / (2 * np.pi), atol=2e-6)
    assert np.allclose(_response(rl, w), np.abs(w) / (4 * np.pi), atol=2e-5)
    d = 0.37
    assert np.allclose(oracle.filter_taps(1, False, d, 50), rl[:50] / d ** 2)
    # equiangular Ram-Lak: h(n d) = -1/(2 pi^2 sin^2(n d)) for odd n (FanProjection.m:241-254 halved)
    eq = oracle.filter_taps(1, True, 0.01, 8)
    assert np.isclose(eq[3], -1 / (2 * np.pi ** 2 * np.sin(0.03) ** 2)) and eq[2] == 0 and np.isclose(eq[0], 1 / (8e-4))


def test_cone_filter_matches_numpy(oracle):
    og = small_cone(oracle, np_=37, nx=5, pp=0.6, px=0.9, nb=3)
    y = np.random.default_rng(4).random(og.nray).astype(np.float32)
    for filt in (0, 1):
        out = oracle.cone_filter(og, y, filt).reshape(3, 5, 37)
        d = float(np.float32(abs(og.p[1] - og.p[0])))
        h = oracle.filter_taps(filt, False, d, 37)
        h2 = np.concatenate([h[:0:-1], h])
        Y = y.reshape(3, 5, 37).astype(np.float64)
        P, X = og.p.astype(np.float64), og.xi.astype(np.float64)
        D = og.s.D
        wgt = D / np.sqrt(D * D + P[None, :] ** 2 + X[:, None] ** 2)
        W = (Y * wgt[None]).astype(np.float32).astype(np.float64)
        ref = np.stack([[d * np.convolve(W[b, x], h2)[36:73] for x in range(5)] for b in range(3)])
        assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max()


def _mag_geometry(og):
    nt, ns, nz = og.s.nt, og.s.ns, og.s.nz
    size = np.array(og.s.size[:])
    r = size / np.array([nt, ns, nz])
    x = (np.arange(nt) + 0.5) * r[0] - size[0] / 2
    y = (np.arange(ns) + 0.5) * r[1] - size[1] / 2
    z = (np.arange(nz) + 0.5) * r[2] - size[2] / 2
    return x, y, z, r


def test_fdk_bp_constant_closed_form(oracle):
    """Constant filtered projections c: every voxel whose sample lies on the detector gets
    c * mag^2 * dbeta per view (FDK.cu:116-120)."""
    og = small_cone(oracle, nt=8, ns=9, nz=7, size=(8.0, 9.0, 7.0), np_=13, nx=11, pp=1.0, px=1.0, D=30.0)
    c = 0.75
    out = oracle.fdk_bp(og, np.full(og.nray, c, np.float32)).reshape(7, 9, 8)
    x, y, z, r = _mag_geometry(og)
    D = og.s.D
    dbeta = abs(float(og.beta[1]) - float(og.beta[0]))
    ref = np.zeros((7, 9, 8))
    for b in og.beta.astype(np.float64):
        u = x[None, :] * np.cos(b) + y[:, None] * np.sin(b)
        s = -x[None, :] * np.sin(b) + y[:, None] * np.cos(b)
        m = D / (D - s)
        P = u * m
        X = z[:, None, None] * m[None]
        ok = (P[None] >= og.p[0]) & (P[None] < og.p[-1]) & (X >= og.xi[0]) & (X < og.xi[-1])
        ref += np.where(ok, c * m[None] ** 2 * dbeta, 0.0)
    assert np.abs(out - ref).max() < 1e-9 * ref.max()


def test_bp_ref_ones_is_sum_of_chords(oracle):
    """All-ones projections: BP = sum over views of the chord of the source-to-centre line
    through the voxel, 2 |d| min_a(h_a/|d_a|) (Backprojection3D.cu:243-288)."""
    og = small_cone(oracle, nt=6, ns=7, nz=5, size=(6.0, 7.7, 4.0), np_=15, nx=13, pp=1.0, px=1.0, D=30.0)
    out = oracle.cone_bp_ref(og, np.ones(og.nray, np.float32)).reshape(5, 7, 6)
    x, y, z, r = _mag_geometry(og)
    D = og.s.D
    ref = np.zeros((5, 7, 6))
    for b in og.beta.astype(np.float64):
        dx = x[None, None, :] + D * np.sin(b)
        dy = y[None, :, None] - D * np.cos(b)
        dz = z[:, None, None] + 0 * dx
        dx, dy = dx + 0 * dz, dy + 0 * dz
        n = np.sqrt(dx ** 2 + dy ** 2 + dz ** 2)
        with np.errstate(divide="ignore"):
            lam = np.minimum(np.minimum(r[0] / 2 / np.abs(dx), r[1] / 2 / np.abs(dy)), r[2] / 2 / np.abs(dz))
        u = x[None, None, :] * np.cos(b) + y[None, :, None] * np.sin(b)
        s = -x[None, None, :] * np.sin(b) + y[None, :, None] * np.cos(b)
        m = D / (D - s)
        ok = (u * m >= og.p[0]) & (u * m < og.p[-1]) & (dz * m >= og.xi[0]) & (dz * m < og.xi[-1])
        ref += np.where(ok, 2 * n * lam, 0.0)
    assert np.abs(out - ref).max() < 1e-9 * ref.max()


def test_fdk_recovers_ball_amplitude(oracle):
    """FDK of a centred uniform ball recovers its amplitude near the centre."""
    n = 40
    og = small_cone(oracle, nt=n, ns=n, nz=n, size=(40.0, 40.0, 40.0), D=200.0, nb=120, np_=96, nx=72,
                    pp=0.5, px=0.5)
    g = (np.arange(n) + 0.5) - 20.0
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    vol = ((X ** 2 + Y ** 2 + Z ** 2) <= 12.0 ** 2).astype(np.float32)
    proj = oracle.cone_fp_exact(og, vol.ravel()).astype(np.float32)
    rec = oracle.fdk_bp(og, oracle.cone_filter(og, proj, 1).astype(np.float32)).reshape(n, n, n)
    core = rec[16:24, 16:24, 16:24]
    assert abs(core.mean() - 1.0) < 0.03, core.mean()
    assert abs(rec[2:6, 2:6, 18:22].mean()) < 0.05


@pytest.mark.parametrize("equiangular", [False, True])
@pytest.mark.parametrize("filt", [0, 1])
def test_fan_fbp_recovers_disk(oracle, equiangular, filt):
    """Flat and equiangular fan FBP (both filters) of a uniform disk recover its amplitude
    inside and zero outside: pins every scale factor (dbeta, pitch, 1/U^2 or 1/L^2, the
    half-filter convention) at once."""
    import ctr
    n = 128
    g0 = np.arange(n) + 0.5 - 64
    Y, X = np.meshgrid(g0, g0, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2)
    disk = (r <= 40).astype(np.float32)
    if equiangular:
        g = ctr.geometry.fan_equiangular(n, 1.0, 360, 256, 1.0, 500.0, 1000.0)
    else:
        g = ctr.geometry.fan_flat(n, 1.0, 360, 256, 1.0, 500.0, 1000.0)
    og = oracle.Fan(n, n, g.size, g.distance, g.beta, g.p)
    proj = oracle.fan_fp_exact(og, disk.ravel()).astype(np.float32)
    f = oracle.fan_filter(og, proj, equiangular, filt).astype(np.float32)
    rec = oracle.fan_fbp_bp(og, f, equiangular).reshape(n, n)
    assert abs(rec[r < 30].mean() - 1.0) < 5e-3
    assert abs(rec[(r > 48) & (r < 60)].mean()) < 5e-3
    # the Shepp-Logan phantom as a whole
    img = ctr.phantoms.shepp_logan_2d(n).astype(np.float32)
    rec2 = oracle.fan_fbp_bp(og, oracle.fan_filter(og, oracle.fan_fp_exact(og, img.ravel()).astype(np.float32),
                                                   equiangular, filt).astype(np.float32), equiangular).reshape(n, n)
    assert np.linalg.norm(rec2 - img) / np.linalg.norm(img) < 0.2


def test_dense_matrix_sirt(oracle):
    """The oracle's matrix-free SIRT equals SIRT3D.m:80-82 / SIRT.m:129-132 run on the explicit
    system matrix, whose columns are the projections of unit vectors."""
    og = small_cone(oracle, nt=5, ns=4, nz=3, size=(5.0, 4.0, 3.0), np_=9, nx=7, pp=0.8, px=0.6, nb=7, D=20.0)
    A = np.stack([oracle.cone_fp_exact(og, np.eye(og.nvox, dtype=np.float32)[j]) for j in range(og.nvox)], 1)
    rng = np.random.default_rng(5)
    xt = rng.random(og.nvox)
    b32 = (A @ xt).astype(np.float32)
    b = b32.astype(np.float64)
    x = np.zeros(og.nvox)
    rows, cols = A.sum(1), A.sum(0)
    res = []
    for k in range(15):
        e = b - A @ x
        res.append(np.linalg.norm(e) / np.linalg.norm(b))
        x = np.maximum(x + 1.0 * (A.T @ (e / (rows + 1e-9))) / (cols + 1e-9), 0.0)
    res.append(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
    xo, r2, r1 = oracle.sirt3d(og, b32, 15, 1.0, 0.0)
    assert np.abs(xo - x).max() < 1e-10
    assert np.allclose(r2, res, rtol=1e-9)
    assert np.all(np.diff(r2) <= 1e-12)
    # L1 form of SIRT3D.m:87
    assert np.isclose(r1[-1], np.abs(b - A @ x).sum() / (b.sum() * og.nvox))


def test_dense_matrix_sirt2d(oracle):
    fg = oracle.Fan(6, 5, (6.0, 5.5), 20.0, (2 * np.pi / 9) * np.arange(1, 10), 0.6 * (np.arange(15) - 7))
    A = np.stack([oracle.fan_fp_exact(fg, np.eye(fg.nvox, dtype=np.float32)[j]) for j in range(fg.nvox)], 1)
    b = (A @ np.random.default_rng(6).random(fg.nvox)).astype(np.float32)
    x = np.zeros(fg.nvox)
    for _ in range(10):
        x = np.clip(x + 2.0 * (A.T @ ((b - A @ x) / (A.sum(1) + 1e-9))) / (A.sum(0) + 1e-9), 0.0, 1.0)
    xo, _, _ = oracle.sirt2d(fg, b, 10, 2.0, 0.0, 1.0)
    assert np.abs(xo - x).max() < 1e-10


def test_grid_modes_agree_on_exact_grids(oracle):
    """The reference's first-pair grid and the end-sample grid coincide on float-exact positions."""
    og = small_cone(oracle, np_=16, nx=12, pp=0.5, px=0.25)
    y = np.random.default_rng(7).random(og.nray).astype(np.float32)
    a = oracle.fdk_bp(og, y)
    with oracle.interp_mode(1):
        b = oracle.fdk_bp(og, y)
    assert np.abs(a - b).max() < 1e-12 * np.abs(a).max()


# ---- parallel beam ----------------------------------------------------------

def small_par(O, nt=7, ns=6, size=(7.0, 6.6), nb=9, np_=11, pitch=0.93):
    theta = np.concatenate([[0.0, np.pi / 2, np.pi / 4], np.random.default_rng(8).uniform(0, 2 * np.pi, nb - 3)])
    t = (np.arange(np_) - (np_ - 1) / 2) * pitch + 0.013
    return O.Par(nt, ns, size, theta, t)


def box_chord_par(g, b, it):
    """Chord of the line (x-Ct) cos + (y-Cs) sin = t through [0,W] x [0,H] (slab method)."""
    th, t = g.theta[b], g.t[it]
    W, H = g.s.size[0], g.s.size[1]
    S = np.array([W / 2 + t * np.cos(th), H / 2 + t * np.sin(th)])
    d = np.array([-np.sin(th), np.cos(th)])
    lo, hi = -np.inf, np.inf
    for a, bmax in enumerate((W, H)):
        if d[a] == 0:
            if not (0 <= S[a] < bmax):
                return 0.0
            continue
        t1, t2 = (0 - S[a]) / d[a], (bmax - S[a]) / d[a]
        lo, hi = max(lo, min(t1, t2)), min(hi, max(t1, t2))
    return max(hi - lo, 0.0)


def test_parallel_siddon_matches_brute_force(oracle):
    g = small_par(oracle)
    img = np.random.default_rng(9).uniform(0, 1, g.nvox).astype(np.float32)
    fp = oracle.par_fp_exact(g, img).reshape(len(g.theta), len(g.t))
    for b in range(len(g.theta)):
        for it in range(len(g.t)):
            assert abs(fp[b, it] - oracle.par_ray_brute(g, img, b, it)) < 1e-9


def test_parallel_fp_of_ones_is_box_chord(oracle):
    g = small_par(oracle)
    fp = oracle.par_fp_exact(g, np.ones(g.nvox, np.float32)).reshape(len(g.theta), len(g.t))
    for b in range(len(g.theta)):
        for it in range(len(g.t)):
            assert abs(fp[b, it] - box_chord_par(g, b, it)) < 1e-9


def test_parallel_fp_of_disk_is_analytic_chord(oracle):
    n, R = 256, 80.0
    x = (np.arange(n) + 0.5) - n / 2
    img = ((x[None, :] ** 2 + x[:, None] ** 2) <= R * R).astype(np.float32)
    theta = np.array([0.1, 0.7, 1.9])
    t = np.linspace(-75.3, 75.3, 41)
    g = oracle.Par(n, n, (float(n), float(n)), theta, t)
    fp = oracle.par_fp_exact(g, img).reshape(3, -1)
    exact = 2 * np.sqrt(np.maximum(R * R - t * t, 0))
    assert np.abs(fp - exact[None, :]).max() / exact.max() < 0.01


@pytest.mark.parametrize("filt", [0, 1])
def test_parallel_fbp_recovers_disk(oracle, filt):
    """Analytic projections of a uniform disk (2 A sqrt(R^2 - t^2)) over 360
    views of 1 degree, half-scaled filter: the FBP interior is A."""
    n, R, A = 128, 40.0, 0.8
    theta = np.deg2rad(np.arange(1, 361))
    t = np.arange(-95.5, 96.0, 1.0)
    g = oracle.Par(n, n, (float(n), float(n)), theta, t)
    proj = np.tile(2 * A * np.sqrt(np.maximum(R * R - t * t, 0)), len(theta)).astype(np.float32)
    filt_out = oracle.par_filter(g, proj, filt=filt)
    img = oracle.par_fbp_bp(g, filt_out.astype(np.float32)).reshape(n, n)
    c = (np.arange(n) + 0.5) - n / 2
    inner = (c[None, :] ** 2 + c[:, None] ** 2) < (0.6 * R) ** 2
    assert abs(img[inner].mean() / A - 1) < 0.02


def test_parallel_adjoint_is_transpose(oracle):
    g = small_par(oracle)
    rng = np.random.default_rng(10)
    x = rng.uniform(0, 1, g.nvox).astype(np.float32)
    y = rng.uniform(0, 1, g.nray).astype(np.float32)
    ax = oracle.par_fp_exact(g, x)
    aty = oracle.par_adjoint(g, y)
    lhs, rhs = float(np.dot(ax, y.astype(np.float64))), float(np.dot(x.astype(np.float64), aty))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # column sums of a one-hot pixel: the pixel's chords over all rays
    e = np.zeros(g.nvox, np.float32)
    e[17] = 1.0
    assert np.isclose(oracle.par_adjoint(g, np.ones(g.nray, np.float32))[17], oracle.par_fp_exact(g, e).sum(),
                      rtol=1e-12)


# ---- curved (equiangular) cone FDK with Parker weights ----------------------

def small_curved(O, start=0.0):
    D, Dsd = 150.0, 300.0
    gamma = 0.002 * (np.arange(241) - 120)
    xi = 1.0 * (np.arange(24) - 11.5)
    fan = abs(gamma[-1] - gamma[0])
    dbeta = 0.01
    beta = start + dbeta * np.arange(int(np.ceil((np.pi + fan) / dbeta)) + 1)
    return O.Curved(48, 48, 8, (48.0, 48.0, 8.0), D, Dsd, beta, gamma, xi, start, fan), (D, Dsd, beta, gamma, xi)


def test_parker_weights_sum_to_one_over_conjugate_rays(oracle):
    """Rays (b, g) and (b + pi + 2g, -g) are one line; over a short scan from 0
    the window (ParkFunction.m) gives them weights summing to 1."""
    fan = 0.5
    rng = np.random.default_rng(17)
    for _ in range(400):
        g = rng.uniform(-fan / 2 + 1e-3, fan / 2 - 1e-3)
        b = rng.uniform(0, np.pi + fan)
        w = oracle.parker(b, 0.0, g, fan)
        bc = b + np.pi + 2 * g if b + np.pi + 2 * g <= np.pi + fan else b - np.pi + 2 * g
        wc = oracle.parker(bc, 0.0, -g, fan) if 0 <= bc <= np.pi + fan else 0.0
        assert 0.0 <= w <= 1.0
        assert abs(w + wc - 1.0) < 1e-12, (b, g, w, wc)
    assert oracle.parker(-0.1, 0.0, 0.0, fan) == 0.0 and oracle.parker(np.pi + fan + 0.1, 0.0, 0.0, fan) == 0.0


def test_curved_taps_match_equiangular_ramp(oracle):
    """Row taps are the SL taps of the arc pitch times (n dg / sin(n dg))^2
    (equiangular ramp, Kak & Slaney 3.4); their DFT is twice the SL-half
    response |sin(w/2)|/(2 pi) of the flat filter at small n dg."""
    g, _ = small_curved(oracle)
    taps = oracle.curved_taps(g, 0)
    P = 300.0 / np.sqrt(300.0 ** 2 + 11.5 ** 2)
    dg = 0.002 * P
    n = np.arange(1, 241)
    flat = -2.0 / (np.pi ** 2 * dg ** 2 * (4 * n ** 2 - 1))
    assert np.allclose(taps[1:], flat * (n * dg / np.sin(n * dg)) ** 2, rtol=1e-12)
    assert np.isclose(taps[0], 2.0 / (np.pi * dg) ** 2, rtol=1e-14)


@pytest.mark.parametrize("start", [0.0, 0.027])
def test_curved_fdk_recovers_ball(oracle, start):
    """Short scan (pi + fan) with Parker weights: the central slices of the
    FDK of an analytic ball are its amplitude (also with the start-angle
    offset that the window's sine terms ignore)."""
    g, (D, Dsd, beta, gamma, xi) = small_curved(oracle, start)
    proj = curved_ball_projections(beta, gamma, xi, D, Dsd, (3.0, -2.0, 0.0), 12.0, 1.0)
    filt = oracle.curved_filter(g, proj)
    with oracle.interp_mode(1):
        vol = oracle.curved_bp(g, filt.astype(np.float32)).reshape(8, 48, 48)
    c = np.arange(48) + 0.5 - 24
    inside = ((c[None, :] - 3.0) ** 2 + (c[:, None] + 2.0) ** 2) < 8.0 ** 2
    outside = ((c[None, :] - 3.0) ** 2 + (c[:, None] + 2.0) ** 2) > 16.0 ** 2
    for k in (3, 4):
        assert abs(vol[k][inside].mean() - 1.0) < 0.02
        assert abs(vol[k][outside].mean()) < 0.02


def test_oracle_row_segments_sum_to_projection(oracle):
    """The per-ray segment lists are the rows behind the exact projectors."""
    g = small_par(oracle)
    img = np.random.default_rng(23).uniform(0, 1, g.nvox).astype(np.float32)
    fp = oracle.par_fp_exact(g, img)
    for ray in range(g.nray):
        idx, ln = oracle.par_segments(g, ray)
        assert len(np.unique(idx)) == len(idx) and np.all(ln > 0)
        assert abs(float(np.dot(ln, img[idx])) - fp[ray]) < 1e-12


# ---- iterative FDK error slice ------------------------------------------------

def small_fdkerr(O, n=16, t=5):
    from ctr.geometry import iterative_fdk_args
    a = iterative_fdk_args(n, n_views=24, pint=1.6)
    return O.FdkErr(n, n, n, a["size"], a["D"], a["beta"], a["u"][::3], a["p"], a["g"], t)


def test_fdk_error_slice_vanishes_for_z_invariant_volumes(oracle):
    """FDK is exact for objects constant along z: every term is a difference
    of two voxels of one (x, y) column."""
    a = small_fdkerr(oracle)
    col = np.random.default_rng(29).random((16, 16)).astype(np.float32)
    vol = np.broadcast_to(col, (16, 16, 16)).copy()
    assert np.abs(oracle.fdk_error_slice(a, vol.ravel())).max() == 0.0


def test_fdk_error_slice_linear_and_z_mirror_symmetric(oracle):
    """Linear in the volume; a volume mirrored about the central z plane gives
    the error mirrored in z (the moved height is odd in z)."""
    a = small_fdkerr(oracle)
    rng = np.random.default_rng(30)
    f = rng.random((16, 16, 16)).astype(np.float32)
    h = rng.random((16, 16, 16)).astype(np.float32)
    ef, eh = oracle.fdk_error_slice(a, f.ravel()), oracle.fdk_error_slice(a, h.ravel())
    e2 = oracle.fdk_error_slice(a, (2 * f + h).ravel())
    assert np.abs(e2 - (2 * ef + eh)).max() < 1e-5 * np.abs(e2).max() and np.abs(ef).max() > 0
    sym = (f + f[::-1]).astype(np.float32)
    e = oracle.fdk_error_slice(a, sym.ravel()).reshape(24, 16, 16)
    assert np.abs(e - e[:, ::-1, :]).max() < 1e-5 * np.abs(e).max()


# ---- bilinear interpolation on affine data ----------------------------------
# Linear interpolation reproduces data that are affine in the sample index
# exactly, so a backprojection of projections R[b][ix][ip] = a_b ip + c_b ix + e_b
# has the closed form sum_b (a_b tp + c_b tx + e_b) * weight_b, with (tp, tx)
# the fractional sample indices of the voxel's detector point.  Per-view
# coefficients and a != c catch a swapped axis, a flipped weight or a wrong
# view; interpolation mode 1 runs on a grid that is not exact in float32,
# mode 0 (the reference's first-pair pitch) on one that is.

def _frac_index(pos32, x):
    """Fractional index of x on the uniform grid through the end samples."""
    p0, pn, n = float(pos32[0]), float(pos32[-1]), len(pos32)
    return (x - p0) / ((pn - p0) / (n - 1))


def _on_detector(pos32, x):
    lo, hi = min(float(pos32[0]), float(pos32[-1])), max(float(pos32[0]), float(pos32[-1]))
    return (x >= lo) & (x < hi)


def _affine_views(rng, nb):
    # multiples of 1/64: the projections are exact in float32
    a, c, e = (rng.integers(-64, 65, nb) / 64.0 for _ in range(3))
    return a, c, e


_BILINEAR_GRIDS = [(1, 0.7, 0.55), (0, 1.0, 0.75)]   # (interp mode, column pitch, row pitch)


def _cone_affine_case(oracle, mode, pp, px, seed):
    og = small_cone(oracle, nt=8, ns=9, nz=7, size=(8.0, 9.9, 6.3), np_=17, nx=13, pp=pp, px=px, D=30.0)
    nb, np_, nx = len(og.beta), len(og.p), len(og.xi)
    a, c, e = _affine_views(np.random.default_rng(seed), nb)
    R = (a[:, None, None] * np.arange(np_)[None, None, :] + c[:, None, None] * np.arange(nx)[None, :, None]
         + e[:, None, None])
    assert np.array_equal(R.astype(np.float32).astype(np.float64), R)
    x, y, z, r = _mag_geometry(og)
    views = []
    for k, b in enumerate(og.beta.astype(np.float64)):
        u = x[None, None, :] * np.cos(b) + y[None, :, None] * np.sin(b)
        w = -x[None, None, :] * np.sin(b) + y[None, :, None] * np.cos(b)
        m = og.s.D / (og.s.D - w)
        P = (u * m) + 0 * z[:, None, None]
        X = z[:, None, None] * m
        ok = _on_detector(og.p, P) & _on_detector(og.xi, X)
        val = a[k] * _frac_index(og.p, P) + c[k] * _frac_index(og.xi, X) + e[k]
        views.append((b, ok, val, m))
    return og, R.astype(np.float32).ravel(), views, (x, y, z, r)


@pytest.mark.parametrize("mode,pp,px", _BILINEAR_GRIDS)
def test_fdk_bp_affine_closed_form(oracle, mode, pp, px):
    """FDK.cu:80-123: bilinear sample x mag^2 x dbeta."""
    og, R, views, _ = _cone_affine_case(oracle, mode, pp, px, 41)
    dbeta = abs(float(og.beta[1]) - float(og.beta[0]))
    ref = sum(np.where(ok, val * m * m * dbeta, 0.0) for _, ok, val, m in views)
    with oracle.interp_mode(mode):
        out = oracle.fdk_bp(og, R).reshape(ref.shape)
    assert np.abs(out - ref).max() < 1e-11 * np.abs(ref).max()
    assert np.count_nonzero(ref) > 0.5 * ref.size


@pytest.mark.parametrize("mode,pp,px", _BILINEAR_GRIDS)
def test_cone_bp_ref_affine_closed_form(oracle, mode, pp, px):
    """Backprojection3D.cu:68-288: bilinear sample x chord of the
    source-to-voxel-centre line through the voxel."""
    og, R, views, (x, y, z, r) = _cone_affine_case(oracle, mode, pp, px, 42)
    D = og.s.D
    ref = 0.0
    for b, ok, val, _ in views:
        dx = x[None, None, :] + D * np.sin(b) + 0 * z[:, None, None] + 0 * y[None, :, None]
        dy = y[None, :, None] - D * np.cos(b) + 0 * dx
        dz = z[:, None, None] + 0 * dx
        with np.errstate(divide="ignore"):
            half = np.minimum(np.minimum(r[0] / 2 / np.abs(dx), r[1] / 2 / np.abs(dy)), r[2] / 2 / np.abs(dz))
        ref = ref + np.where(ok, val * 2 * half * np.sqrt(dx * dx + dy * dy + dz * dz), 0.0)
    with oracle.interp_mode(mode):
        out = oracle.cone_bp_ref(og, R).reshape(ref.shape)
    assert np.abs(out - ref).max() < 1e-11 * np.abs(ref).max()


def _fan_case(oracle, pp, equiangular, nb=13):
    import ctr
    D, n = 40.0, 11
    beta = (2 * np.pi / nb) * np.arange(1, nb + 1)
    if equiangular:
        g = ctr.geometry.fan_equiangular(n, 1.1, nb, 27, 0.9, D, 2 * D)
        p = g.p
    else:
        p = pp * (np.arange(27) - 13.0)
    og = oracle.Fan(n, n - 2, (n * 1.1, (n - 2) * 0.95), D, beta, p)
    x = (np.arange(n) + 0.5) * 1.1 - n * 1.1 / 2
    y = (np.arange(n - 2) + 0.5) * 0.95 - (n - 2) * 0.95 / 2
    return og, x[None, :], y[:, None]


def _fan_affine(rng, og):
    a, _, e = _affine_views(rng, len(og.beta))
    R = a[:, None] * np.arange(len(og.p))[None, :] + e[:, None]
    return a, e, R.astype(np.float32).ravel()


@pytest.mark.parametrize("mode,pp", [(1, 0.7), (0, 0.5)])
def test_fan_fbp_flat_affine_closed_form(oracle, mode, pp):
    """FBPfan.cu:45-91: linear sample at p/U, weighted 1/U^2 x dbeta (float angles)."""
    og, x, y = _fan_case(oracle, pp, False)
    a, e, R = _fan_affine(np.random.default_rng(43), og)
    p32 = og.p.astype(np.float32)
    b32 = og.beta.astype(np.float32).astype(np.float64)
    D, dbeta = og.s.D, abs(b32[1] - b32[0])
    ref = 0.0
    for k, b in enumerate(b32):
        u = x * np.cos(b) + y * np.sin(b)
        w = -x * np.sin(b) + y * np.cos(b)
        U = (D - w) / D
        ok = _on_detector(p32, u / U)
        ref = ref + np.where(ok, (a[k] * _frac_index(p32, u / U) + e[k]) / (U * U) * dbeta, 0.0)
    with oracle.interp_mode(mode):
        out = oracle.fan_fbp_bp(og, R, False).reshape(ref.shape)
    assert np.abs(out - ref).max() < 1e-11 * np.abs(ref).max()


def test_fan_fbp_equiangular_affine_closed_form(oracle):
    """FanProjection.m:299-324: linear sample at gamma = atan2(u, |w - D|) on the
    channel-angle grid, weighted 1/L^2, times dbeta."""
    og, x, y = _fan_case(oracle, None, True)
    a, e, R = _fan_affine(np.random.default_rng(44), og)
    gam = np.arctan(og.p / og.s.D)            # channel angles of the double positions
    D, dbeta = og.s.D, abs(og.beta[1] - og.beta[0])
    ref = 0.0
    for k, b in enumerate(og.beta):
        u = x * np.cos(b) + y * np.sin(b)
        w = -x * np.sin(b) + y * np.cos(b)
        gb = np.arctan2(u, np.abs(w - D))
        ok = _on_detector(gam, gb)
        ref = ref + np.where(ok, (a[k] * _frac_index(gam, gb) + e[k]) / (u * u + (w - D) ** 2), 0.0)
    ref = ref * dbeta
    with oracle.interp_mode(1):
        out = oracle.fan_fbp_bp(og, R, True).reshape(ref.shape)
    assert np.abs(out - ref).max() < 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("mode,pp", [(1, 0.7), (0, 0.5)])
def test_fan_bp_ref_affine_closed_form(oracle, mode, pp):
    """BackprojectionFan2D.cu:36-113 (entry face corrected): linear sample at
    p/U times the chord of the source-to-pixel-centre line through the pixel."""
    og, x, y = _fan_case(oracle, pp, False)
    a, e, R = _fan_affine(np.random.default_rng(45), og)
    p32 = og.p.astype(np.float32)
    D = og.s.D
    rt, rs = og.s.size[0] / og.s.nt, og.s.size[1] / og.s.ns
    ref = 0.0
    for k, b in enumerate(og.beta.astype(np.float32).astype(np.float64)):
        u = x * np.cos(b) + y * np.sin(b)
        w = -x * np.sin(b) + y * np.cos(b)
        Pd = u / ((D - w) / D)
        dx, dy = x + D * np.sin(b) + 0 * y, y - D * np.cos(b) + 0 * x
        with np.errstate(divide="ignore"):
            chord = np.sqrt(dx * dx + dy * dy) * np.minimum(rt / np.abs(dx), rs / np.abs(dy))
        ok = _on_detector(p32, Pd)
        ref = ref + np.where(ok, (a[k] * _frac_index(p32, Pd) + e[k]) * chord, 0.0)
    with oracle.interp_mode(mode):
        out = oracle.fan_bp_ref(og, R).reshape(ref.shape)
    assert np.abs(out - ref).max() < 1e-11 * np.abs(ref).max()
    assert np.count_nonzero(ref) == ref.size


def test_fan_bp_ref_quirk_chord_to_far_end(oracle):
    """CTR_REF_QUIRKS (BackprojectionFan2D.cu:54-109): the entry point is the
    crossing of the source-to-centre line with the pixel's t-low edge when
    that lies in [0, s_high], else with its s-low edge when that lies in
    [0, t_high]; the weight is 2 |entry - F|, F the ray's end point 2D from
    the source.  Ones in, so the sum over views of those lengths comes out."""
    og, x, y = _fan_case(oracle, 0.5, False)
    D = og.s.D
    nt, ns = og.s.nt, og.s.ns
    rt, rs = og.s.size[0] / nt, og.s.size[1] / ns
    Ct, Cs = og.s.size[0] / 2, og.s.size[1] / 2
    p32 = og.p.astype(np.float32)
    T = np.arange(nt)[None, :] + 0 * y
    S = np.arange(ns)[:, None] + 0 * x
    ref = np.zeros((ns, nt))
    for b in og.beta.astype(np.float32).astype(np.float64):
        u = x * np.cos(b) + y * np.sin(b)
        w = -x * np.sin(b) + y * np.cos(b)
        Pd = u / ((D - w) / D)
        src = np.array([Ct - D * np.sin(b), Cs + D * np.cos(b)])
        th = np.arctan(Pd / D)
        F = (Ct + 2 * D * np.sin(th) * np.cos(b) - (D - 2 * D * np.cos(th)) * np.sin(b),
             Cs + 2 * D * np.sin(th) * np.sin(b) + (D - 2 * D * np.cos(th)) * np.cos(b))
        # the source-to-F line crosses t = T rt at s = src_s + (T rt - src_t) (F_s - src_s)/(F_t - src_t)
        tl = T * rt
        s_at = src[1] + (tl - src[0]) * (F[1] - src[1]) / (F[0] - src[0])
        sl = S * rs
        t_at = src[0] + (sl - src[1]) * (F[0] - src[0]) / (F[1] - src[1])
        use_t = (s_at >= 0) & (s_at <= (S + 1) * rs)
        use_s = ~use_t & (t_at >= 0) & (t_at <= (T + 1) * rt)
        et = np.where(use_t, tl, t_at)
        es = np.where(use_t, s_at, sl)
        L = 2 * np.sqrt((et - F[0]) ** 2 + (es - F[1]) ** 2)
        ok = _on_detector(p32, Pd) & (use_t | use_s)
        ref += np.where(ok, L, 0.0)
    out = oracle.fan_bp_ref(og, np.ones(og.nray, np.float32), quirks=True).reshape(ns, nt)
    assert np.abs(out - ref).max() < 1e-9 * np.abs(ref).max()


# ---- the error slice's weight factor ------------------------------------------

def test_fdk_error_slice_weight_for_volume_linear_in_z(oracle):
    """IterativeFDKerrCompute.cu:66-70 weights each difference of two voxels
    by 0.5 LR^2 sqrt(1 + (P/D)^2) G[f] D/sqrt(D^2 + P^2) dU dP dbeta, and the
    two square-root factors cancel.  For a volume k (16 + 2 j + i)/16 (k, j, i
    the z, s, t voxel indices) every difference is exact in float32,
    (k' - k) (16 + 2 j' + i')/16, so with random (asymmetric) taps G the
    sum is fixed by the voxel indices the kernel looks up (restated from
    :37-58 in float32, as the kernel computes them) and the weight
    0.5 LR^2 G[f] dU dP dbeta.  A dropped or extra factor, a wrong power of
    LR or a mirrored tap index fails it."""
    n, t = 16, 5
    from ctr.geometry import iterative_fdk_args
    A = iterative_fdk_args(n, n_views=24, pint=1.6)
    u = A["u"][::3]
    rng = np.random.default_rng(46)
    G = rng.standard_normal(len(A["g"]))
    a = oracle.FdkErr(n, n, n, A["size"], A["D"], A["beta"], u, A["p"], G, t)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    vol = (k * (16 + 2 * j + i) / 16.0).astype(np.float32)
    out = oracle.fdk_error_slice(a, vol.ravel()).reshape(len(A["beta"]), n, n)

    f32 = np.float32
    D, size = A["D"], A["size"]
    r = [size[q] / n for q in range(3)]
    C = [size[q] / 2 for q in range(3)]
    P = A["p"]
    PInt = abs(P[1] - P[0])
    dbeta = abs(f32(A["beta"][1]) - f32(A["beta"][0]))
    dU = abs(f32(u[1]) - f32(u[0]))
    ref = np.zeros_like(out)
    for ib, beta in enumerate(A["beta"]):
        cb, sb = f32(np.cos(np.float64(beta))), f32(np.sin(np.float64(beta)))
        it = f32((t + 0.5) * r[0] - C[0])
        for z in range(n):
            iz = f32((z + 0.5) * r[2] - C[2])
            for s in range(n):
                is_ = f32((s + 0.5) * r[1] - C[1])
                dt = f32(it * cb) + f32(is_ * sb)
                ds = f32(is_ * cb) - f32(it * sb)
                LR = f32(D / (D - np.float64(ds)))
                dtl = np.float64(f32(dt * LR))
                ff = np.floor((dtl - P) / PInt) + len(P)                       # [np]
                for U in u:
                    U = f32(U)
                    num = f32(f32(ds - U) * iz)
                    Zd = f32(np.float64(iz) + np.float64(num) / (D - np.float64(ds)))
                    T = (P * (D - np.float64(U)) / D).astype(f32)
                    X = f32(T * cb) - f32(U * sb)
                    Y = f32(T * sb) + f32(U * cb)
                    fx = np.floor((X.astype(np.float64) + C[0]) / r[0])
                    fy = np.floor((Y.astype(np.float64) + C[1]) / r[1])
                    fz = np.floor((np.float64(Zd) + C[2]) / r[2])
                    ok = (ff >= 0) & (ff < len(G)) & (fx >= 0) & (fx < n) & (fy >= 0) & (fy < n) & (fz >= 0) & (fz < n)
                    diff = (fz - z) * (16 + 2 * fy + fx) / 16.0
                    w = 0.5 * np.float64(LR) ** 2 * G[np.where(ok, ff, 0).astype(int)] * dU * PInt * dbeta
                    ref[ib, z, s] += np.sum(np.where(ok, diff * w, 0.0))
    assert np.abs(ref).max() > 0
    assert np.abs(out - ref).max() < 1e-9 * np.abs(ref).max()
