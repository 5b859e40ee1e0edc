"""Closed-form line integrals used as test inputs (test helpers, CPU)."""
import numpy as np


def curved_ball_projections(beta, gamma, xi, D, Dsd, centre, radius, amp):
    """Line integrals of a uniform ball on the cylindrical detector: the ray of
    (beta, gamma, xi) leaves S = (-D sin b, D cos b, 0) along
    (sin(b+g), -cos(b+g), xi/Dsd) (so that gamma = atan(u/(D-s)) and
    xi = Dsd z/rho for the points on it).  Chord of a ball: 2 sqrt(r^2 - d^2),
    d the distance of the centre from the line.  Layout [beta][xi][gamma],
    float32, computed one view at a time."""
    g = np.asarray(gamma, np.float64)[None, :]
    x = np.asarray(xi, np.float64)[:, None]
    c = np.asarray(centre, np.float64)
    out = np.empty((len(beta), len(xi), len(gamma)), np.float32)
    for k, b in enumerate(np.asarray(beta, np.float64)):
        w = np.array([-D * np.sin(b), D * np.cos(b), 0.0]) - c
        dx, dy = np.sin(b + g) + 0 * x, -np.cos(b + g) + 0 * x
        dz = x / Dsd + 0 * g
        n = np.sqrt(dx * dx + dy * dy + dz * dz)
        proj = (w[0] * dx + w[1] * dy + w[2] * dz) / n
        dist2 = w @ w - proj * proj
        out[k] = 2 * amp * np.sqrt(np.maximum(radius * radius - dist2, 0.0))
    return out.ravel()


def iterative_fdk_args(n, n_views=360, size=60.0, D=200.0, pint=0.4):
    """The configuration of IterativeFDK.m:5-55,293-312 for an n^3 volume over
    size mm: P = -1.1 Rplane : pint : 1.1 Rplane, U = Rpic - dU (1..LU) with
    dU the voxel pitch, views 1..360 degrees, half-less SL taps G [2 LP - 1]."""
    res = size / n
    rpic, rplane = size * np.sqrt(3) / 2, size * np.sqrt(2) / 2
    maxp = rplane * 1.1
    p = np.arange(-maxp, maxp + 1e-9, pint)
    lp = len(p)
    du = res
    lu = int(np.floor(2 * rpic / du))
    u = (rpic - du * np.arange(1, lu + 1)).astype(np.float32)
    beta = np.deg2rad(np.arange(1, n_views + 1) * (360.0 / n_views)).astype(np.float32)
    k = np.arange(lp)
    half = -1.0 / (np.pi * pint) ** 2 / (4.0 * k ** 2 - 1.0)
    half[0] = 1.0 / (np.pi * pint) ** 2
    g = np.concatenate([half[:0:-1], half])
    return dict(size=(size, size, size), D=D, beta=beta, u=u, p=p, g=g)
