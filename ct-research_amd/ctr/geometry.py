"""Physical scanner configurations mapped onto the reference's arguments.

The reference's detector is a virtual flat detector through the isocentre:
a physical flat-panel pitch maps to pitch*SOD/SDD there, and an equiangular
fan bin at angle gamma maps to p = SOD*tan(gamma).  View angles start at
dbeta, not 0 (Analytic_Algorithm/FBP_type_analytic_algorithm_GPU/FDK_GPU.m:63-67).
"""
import numpy as np

from ._lib import ConeGeom, FanGeom


def centered_positions(n, pitch):
    """pitch * ((0:n-1) - (n-1)/2), the symmetric sample grid of FDK_GPU.m:47-58."""
    return pitch * (np.arange(n, dtype=np.float64) - (n - 1) / 2.0)


def views(n, total=2 * np.pi):
    """beta = d:d:total with d = total/n (FDK_GPU.m:63-67)."""
    d = total / n
    return d * np.arange(1, n + 1, dtype=np.float64)


def cone_circular(n_vox, voxel_mm, n_views, det_u, det_v, det_pitch, sod, sdd, nz=None):
    """Circular cone-beam geometry from physical parameters."""
    nz = n_vox if nz is None else nz
    iso = det_pitch * sod / sdd
    size = (n_vox * voxel_mm, n_vox * voxel_mm, nz * voxel_mm)
    return ConeGeom(n_vox, n_vox, nz, size, sod, views(n_views).astype(np.float32),
                    centered_positions(det_u, iso).astype(np.float32),
                    centered_positions(det_v, iso).astype(np.float32))


def fan_equiangular(n_pix, pixel_mm, n_views, n_bins, bin_pitch, sod, sdd):
    """Fan beam with an equiangular detector: gamma pitch = bin pitch / SDD."""
    dg = bin_pitch / sdd
    gam = dg * (np.arange(n_bins) - (n_bins - 1) / 2.0)
    return FanGeom(n_pix, n_pix, (n_pix * pixel_mm, n_pix * pixel_mm), sod, views(n_views), sod * np.tan(gam))


def fan_flat(n_pix, pixel_mm, n_views, n_bins, bin_pitch, sod, sdd):
    iso = bin_pitch * sod / sdd
    return FanGeom(n_pix, n_pix, (n_pix * pixel_mm, n_pix * pixel_mm), sod, views(n_views),
                   centered_positions(n_bins, iso))


# Workloads of the benchmark configurations.
def config_fan_2d():
    """2-D fan beam: 256^2 Shepp-Logan at 1 mm, 360 views, 512 equiangular bins of 1 mm at SDD 1000, SOD 500."""
    return fan_equiangular(256, 1.0, 360, 512, 1.0, 500.0, 1000.0)


def config_cone_256(n_views=360):
    """3-D cone beam: 256^3 at 1 mm, n_views over 360 deg, 512x512 flat detector at 1 mm, SOD 600, SDD 1200."""
    return cone_circular(256, 1.0, n_views, 512, 512, 1.0, 600.0, 1200.0)


def config_cone_512():
    """512^3 at 0.5 mm, 720 views, 1024x768 detector at 0.8 mm, SOD 700, SDD 1100."""
    return cone_circular(512, 0.5, 720, 1024, 768, 0.8, 700.0, 1100.0)


def iterative_fdk_args(n, n_views=360, size=60.0, D=200.0, pint=0.4):
    """The configuration of IterativeFDK.m:5-55,293-312 (My_novel_algorithm/IterativeFDK) for an n^3 volume over
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
