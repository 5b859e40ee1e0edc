"""Input generators: the reference phantoms in the projector's layout, the
seeded ellipsoid generator and the layout helpers.  CPU only."""
import numpy as np


def test_phantom3d_values_and_volume():
    import ctr
    n = 64
    v = ctr.phantoms.phantom3d(n, "shepp-logan")
    assert v.shape == (n, n, n) and v.dtype == np.float32
    # the outer shell: amplitude 2 between the two outer ellipsoids, 2 - .98 inside
    e = ctr.phantoms.ellipsoids_3d("shepp-logan")
    assert np.isclose(v.max(), 2.0) and v.min() == 0.0
    # voxel count of the outer ellipsoid = its volume / voxel volume (grid step 2/(n-1))
    inside = (ctr.phantoms.rasterize_3d(e[:1] * np.array([1, 1, 1, 1, 1, 1, 1, 1]), n) > 0).sum()
    vol = 4 / 3 * np.pi * e[0, 1] * e[0, 2] * e[0, 3] / (2 / (n - 1)) ** 3
    assert abs(inside / vol - 1) < 0.02


def test_phantom3d_axis_convention():
    """MATLAB's p(i,j,k) has x = rng(j), y = rng(i); reshape(p,[],1) puts i fastest, so the
    projector's t axis is MATLAB y and s is x (FDK_GPU.m:74).  Ellipsoid 6 of phantom3d.m is
    centred at (x,y,z) = (0, .35, -.25) with semi-axes (.21, .25, .5)."""
    import ctr
    n = 65
    e = ctr.phantoms.ellipsoids_3d("modified shepp-logan")[5:6]
    v = ctr.phantoms.rasterize_3d(e, n)
    r = (np.arange(n) - 32) / 32.0
    k = int(np.argmin(np.abs(r + 0.25)))
    i_y = int(np.argmin(np.abs(r - 0.35)))
    j_x = int(np.argmin(np.abs(r - 0.0)))
    assert np.isclose(v[k, j_x, i_y], e[0, 0])     # [z][s = x][t = y]
    assert v[k, i_y, j_x] == 0.0                   # swapped axes would land outside


def test_shepp_logan_2d_orientation():
    """img[s][t] = f(x = xax[t], y = xax[s]): ellipse 5 (0.01 at (0, .35)) sits in the upper half."""
    import ctr
    n = 129
    img = ctr.phantoms.shepp_logan_2d(n)
    r = (np.arange(n) - 64) / 64.0
    s_up, s_dn, t0 = int(np.argmin(np.abs(r - 0.35))), int(np.argmin(np.abs(r + 0.35))), 64
    assert np.isclose(img[s_up, t0], 1 - 0.98 + 0.01)
    assert np.isclose(img[s_dn, t0], 1 - 0.98)
    # and agrees with the MATLAB layout through Img2vec_Mat2Cpp2D
    mat = np.flipud(img)                   # MATLAB row 1 = top (y = +1)
    assert np.array_equal(ctr.metrics.img2vec(mat), img.ravel().astype(np.float32))
    assert np.array_equal(ctr.metrics.vec2img(ctr.metrics.img2vec(mat), n, n), mat.astype(np.float32))


def test_xoshiro_known_answers():
    import ctr
    X = ctr.phantoms.Xoshiro256ss
    g = X(0)
    # splitmix64(0) first output seeds s[0]
    assert g.s[0] == 0xE220A8397B1DCDAF
    g.s = [1, 2, 3, 4]
    assert g.next() == 11520          # rotl(2*5, 7) * 9
    assert g.next() == 0
    a = [X(20261202).uniform() for _ in range(1)] + [X(20261202).uniform()]
    assert a[0] == a[1] and 0.0 <= a[0] < 1.0


def test_random_ellipsoids_ranges():
    import ctr
    for seed in range(20261203, 20261209):
        e = ctr.phantoms.random_ellipsoids(seed)
        assert 10 <= len(e) <= 40
        assert np.all((e[:, 1:4] >= 0.1 - 1e-12) & (e[:, 1:4] <= 0.8 + 1e-12))
        assert np.all((e[:, 4:7] ** 2).sum(1) <= 0.25 + 1e-12)
        assert np.all((e[:, 0] >= 0) & (e[:, 0] <= 1))
    assert np.array_equal(ctr.phantoms.random_ellipsoids(7), ctr.phantoms.random_ellipsoids(7))


def test_poisson_noise():
    import ctr
    b = np.full(200000, 2.0, np.float32)
    nb = ctr.phantoms.poisson_noise(b, 1e5, 3)
    # -ln(N/I0) with N ~ Poisson(I0 e^-b): mean b, variance ~ 1/(I0 e^-b)
    assert abs(nb.mean() - 2.0) < 1e-3
    assert abs(nb.var() * 1e5 * np.exp(-2.0) - 1) < 0.05
    assert np.array_equal(nb, ctr.phantoms.poisson_noise(b, 1e5, 3))


def test_metrics():
    import ctr
    x = np.array([1.0, 2.0, 3.0])
    assert ctr.metrics.rrmse(x, x) == 0.0
    assert np.isclose(ctr.metrics.rrmse(2 * x, x), 1.0, rtol=1e-12)
    assert np.isclose(ctr.metrics.local_error(x + 1, x), np.sqrt(3) / np.sqrt(14))


def test_torch_rasterizer_matches_numpy():
    from ctr import phantoms
    e = phantoms.random_ellipsoids(20261202)
    a = phantoms.rasterize_3d(e, 24)
    b = phantoms.rasterize_3d_torch(e, 24, device="cpu").numpy()
    # same masks; the float64 amplitude sums may round differently in the last float32 bit
    assert np.abs(a - b).max() <= 1e-6 and np.array_equal(a > 0, b > 0)
