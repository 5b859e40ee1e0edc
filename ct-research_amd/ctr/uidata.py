"""Real-scanner short-scan data in the layout of
My_toolbox/Implement_UnitedImaging_Data/ImplementUIData.m and
Analytic_Algorithm/FBP_type_analytic_algorithm/FanProjection.m:31-43.

On disk:
  detectorangle.dat  raw float32 [n_gamma] channel angles (rad)
  data.dat           raw float32 [n_views][n_rows][n_gamma] line integrals, channel fastest

The scan of ImplementUIData.m:7-39: 936 channels at 9.7527e-4 rad, 80 rows
at 0.915 mm on the detector, 800 views over 240 degrees starting at
1.5466 degrees, SOD 570 mm, SDD 1060.2 mm, a 512 x 512 x 50 volume over
512 x 512 x 23 mm.  The script negates the stored channel angles
(ImplementUIData.m:20) and takes the fan angle as |gamma_last - gamma_first|.
"""
import numpy as np

from ._lib import CurvedGeom

N_GAMMA, N_ROWS, N_VIEWS = 936, 80, 800
GAMMA_PITCH = 9.7527e-4
ROW_PITCH = 0.9150
SOD, SDD = 570.0, 1060.2
SCAN_ARC_DEG, START_DEG = 240.0, 1.5466


def load_detector_angles(path, n_gamma=N_GAMMA):
    """detectorangle.dat: n_gamma float32 angles (FanProjection.m:31-34)."""
    a = np.fromfile(path, dtype="<f4", count=n_gamma)
    if a.size != n_gamma:
        raise ValueError(f"{path}: expected {n_gamma} float32 angles, found {a.size}")
    return a


def load_projections(path, n_gamma=N_GAMMA, n_rows=N_ROWS, n_views=N_VIEWS):
    """data.dat: float32 [n_views][n_rows][n_gamma] (ImplementUIData.m:41-44), memory-mapped."""
    m = np.memmap(path, dtype="<f4", mode="r")
    need = n_gamma * n_rows * n_views
    if m.size < need:
        raise ValueError(f"{path}: expected {need} float32 samples, found {m.size}")
    return m[:need].reshape(n_views, n_rows, n_gamma)


def synthetic_detector_angles(n_gamma=N_GAMMA, pitch=GAMMA_PITCH):
    """Uniform channel angles centred on the central ray, in the stored sign
    convention (the scan negates them)."""
    return (pitch * (np.arange(n_gamma) - (n_gamma - 1) / 2)).astype(np.float32)


def scan_geometry(detector_angles, n_views=N_VIEWS, nt=512, ns=512, nz=50, size=(512.0, 512.0, 23.0)):
    """The CurvedGeom of ImplementUIData.m:7-39 for the given stored channel angles."""
    gamma = -np.asarray(detector_angles, np.float64)
    n_rows = N_ROWS
    max_xi = (n_rows - 1) / 2 * ROW_PITCH
    xi = np.linspace(-max_xi, max_xi, n_rows)
    dbeta = np.deg2rad(SCAN_ARC_DEG / n_views)
    start = np.deg2rad(START_DEG)
    beta = start + dbeta * np.arange(n_views)
    fan = abs(gamma[-1] - gamma[0])
    return CurvedGeom(nt, ns, nz, size, SOD, SDD, beta, gamma, xi, start, fan)
