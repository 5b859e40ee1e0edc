"""Test inputs: the reference's analytic phantoms, seeded random ellipsoids
and Poisson noise.  All volumes are returned in the C layout of the
projector ([z][s][t], t fastest) and images as [s][t] (t = x to the right,
s = y upward), i.e. already converted as Img2vec_Mat2Cpp2D.m:2 (2-D) and a
plain reshape (3-D, FDK_GPU.m:74) would convert the MATLAB arrays.
"""
import numpy as np

# My_toolbox/My_phantom/phantom3d.m:118-137 (modified Shepp-Logan; the
# 'Shepp-Logan' variant replaces column A, :110-113).
#            A      a      b      c      x0      y0      z0    phi
_MOD_SL_3D = np.array([
    [1.0, .6900, .920, .900, 0, 0, 0, 0],
    [-.8, .6624, .874, .880, 0, 0, 0, 0],
    [-.2, .310, .1100, .220, .22, 0, -.25, 72],
    [-.2, .410, .1600, .280, -.22, 0, -.25, 108],
    [.2, .0460, .046, .046, 0, .1, -.25, 0],
    [.2, .2100, .250, .500, 0, .35, -.25, 0],
    [.1, .0460, .046, .046, 0, -.1, -.25, 0],
    [.1, .0460, .023, .020, -.08, -.605, -.25, 90],
    [.1, .0230, .023, .020, 0, -.606, -.25, 90],
    [.1, .0460, .023, .020, .06, -.605, -.25, 0],
    [.2, .0560, .040, .10, .06, -.105, .625, 0],
    [-.2, .0560, .056, .10, 0, .105, .625, 0],
])
_SL_3D_AMPL = np.array([2, -.98, -.02, -.02, .02, .02, .01, .01, .01, .01, .02, -.02])

# My_toolbox/My_phantom/StandardPhantom.m:74-89 ('shepp-logan', values 0..1)
#              A      a      b     x0     y0    phi
_SL_2D = np.array([
    [1, .69, .92, 0, 0, 0],
    [-.98, .6624, .8740, 0, -.0184, 0],
    [-.02, .1100, .3100, .22, 0, -18],
    [-.02, .1600, .4100, -.22, 0, 18],
    [.01, .2100, .2500, 0, .35, 0],
    [.01, .0460, .0460, 0, .1, 0],
    [.01, .0460, .0460, 0, -.1, 0],
    [.01, .0460, .0230, -.08, -.605, 0],
    [.01, .0230, .0230, 0, -.606, 0],
    [.01, .0230, .0460, .06, -.605, 0],
])


def ellipsoids_3d(kind="shepp-logan", n_ellipsoids=12):
    """Ellipsoid table of phantom3d.m; n_ellipsoids=10 drops the two z0=.625 rows (:135-136)."""
    e = _MOD_SL_3D.copy()
    if kind == "shepp-logan":
        e[:, 0] = _SL_3D_AMPL
    elif kind != "modified shepp-logan":
        raise ValueError(kind)
    return e[:n_ellipsoids]


def _axis(n):
    return (np.arange(n) - (n - 1) / 2.0) / ((n - 1) / 2.0)


def rasterize_3d(e, n, dtype=np.float32):
    """Sum of ellipsoid indicators at the voxel centres of an n^3 grid on
    [-1,1]^3 (phantom3d.m:71-110).  MATLAB's p(i,j,k) has x = rng(j),
    y = rng(i), z = rng(k); its column-major reshape puts i (y) fastest, so
    the C layout [z][s][t] has t <-> y and s <-> x."""
    r = _axis(n)
    y = r[None, None, :]   # t axis
    x = r[None, :, None]   # s axis
    out = np.zeros((n, n, n), dtype=np.float64)
    for k in range(len(e)):
        A, a, b, c, x0, y0, z0, phi = e[k]
        ph = np.deg2rad(phi)
        cp, sp = np.cos(ph), np.sin(ph)
        # rotated coordinates (phantom3d.m:97-104): alpha = [c s 0; -s c 0; 0 0 1]
        sx, sy = cp * x0 + sp * y0, -sp * x0 + cp * y0
        u = (cp * x + sp * y - sx) ** 2 / a ** 2 + (-sp * x + cp * y - sy) ** 2 / b ** 2   # (1, n(s), n(t))
        for iz in range(n):
            w = (r[iz] - z0) ** 2 / c ** 2
            m = (u + w) <= 1.0
            if m.any():
                out[iz][m[0]] += A
    return out.astype(dtype)


def rasterize_3d_torch(e, n, device="cuda"):
    """rasterize_3d on a torch device (same float64 arithmetic, same layout);
    returns a float32 tensor [z][s][t]."""
    import torch
    r = torch.from_numpy(_axis(n)).to(device)
    y = r[None, None, :]   # t axis
    x = r[None, :, None]   # s axis
    z = r[:, None, None]
    out = torch.zeros((n, n, n), dtype=torch.float64, device=device)
    for k in range(len(e)):
        A, a, b, c, x0, y0, z0, phi = (float(v) for v in e[k])
        ph = np.deg2rad(phi)
        cp, sp = float(np.cos(ph)), float(np.sin(ph))
        sx, sy = cp * x0 + sp * y0, -sp * x0 + cp * y0
        u = (cp * x + sp * y - sx) ** 2 / a ** 2 + (-sp * x + cp * y - sy) ** 2 / b ** 2
        w = (z - z0) ** 2 / c ** 2
        out += A * ((u + w) <= 1.0)
    return out.to(torch.float32)


def phantom3d(n, kind="shepp-logan", n_ellipsoids=12, dtype=np.float32):
    return rasterize_3d(ellipsoids_3d(kind, n_ellipsoids), n, dtype)


def shepp_logan_2d(n, dtype=np.float64):
    """StandardPhantom('shepp-logan', n) in the C layout img[s][t]
    (StandardPhantom.m:68-88 then Img2vec_Mat2Cpp2D.m:2): img[s][t] = f(x=xax[t], y=xax[s])."""
    r = _axis(n)
    x = r[None, :]
    y = r[:, None]
    out = np.zeros((n, n))
    for A, a, b, x0, y0, phi in _SL_2D:
        ph = np.deg2rad(phi)
        cp, sp = np.cos(ph), np.sin(ph)
        xx, yy = x - x0, y - y0
        m = ((xx * cp + yy * sp) ** 2 / a ** 2 + (yy * cp - xx * sp) ** 2 / b ** 2) <= 1.0
        out[m] += A
    return out.astype(dtype)


class Xoshiro256ss:
    """xoshiro256** (Blackman & Vigna 2018), seeded through splitmix64, with
    uniform doubles from the top 53 bits: reproducible on any platform."""

    M = (1 << 64) - 1

    def __init__(self, seed):
        s = seed & self.M
        self.s = []
        for _ in range(4):
            s = (s + 0x9E3779B97F4A7C15) & self.M
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.M
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.M
            self.s.append(z ^ (z >> 31))

    @staticmethod
    def _rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & Xoshiro256ss.M

    def next(self):
        s = self.s
        result = (self._rotl((s[1] * 5) & self.M, 7) * 9) & self.M
        t = (s[1] << 17) & self.M
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return result

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def randint(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1)) if hi > lo else lo


def random_ellipsoids(seed, n_min=10, n_max=40, axis_lo=0.05, axis_hi=0.40):
    """Seeded random ellipsoid table in phantom3d units (FOV = [-1,1]^3, so a
    semi-axis of f*FOV is 2f): 10-40 ellipsoids, semi-axes uniform in
    [axis_lo, axis_hi]*FOV, centres uniform in the ball of radius 0.5*FOV/2,
    rotation about z uniform, amplitude uniform in [0, 1]."""
    g = Xoshiro256ss(seed)
    m = g.randint(n_min, n_max)
    rows = []
    for _ in range(m):
        a, b, c = (2 * g.uniform(axis_lo, axis_hi) for _ in range(3))
        while True:
            cx, cy, cz = (g.uniform(-0.5, 0.5) for _ in range(3))
            if cx * cx + cy * cy + cz * cz <= 0.25:
                break
        phi = g.uniform(0.0, 180.0)
        A = g.uniform(0.0, 1.0)
        rows.append([A, a, b, c, cx, cy, cz, phi])
    return np.array(rows)


def random_phantom(n, seed, dtype=np.float32):
    return rasterize_3d(random_ellipsoids(seed), n, dtype)


def poisson_noise(b, i0, seed):
    """b_noisy = -ln(max(Poisson(I0 exp(-b)), 1) / I0), seeded."""
    rng = np.random.default_rng(seed)
    counts = rng.poisson(i0 * np.exp(-np.asarray(b, dtype=np.float64)))
    return (-np.log(np.maximum(counts, 1) / i0)).astype(np.float32)
