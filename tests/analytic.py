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

