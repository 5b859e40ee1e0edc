"""The reference's MEX operators with their MATLAB call signatures.

Each function takes the arguments in the order and meaning of the MEX
gateway it replaces and returns a float64 column vector, as
mxCreateDoubleMatrix(L,1) does there.  Differences from the reference,
all deliberate: invalid arguments raise (the reference printed a message
and returned zeros, e.g. ProjectionCone_3D.cpp:38-42); the device context
persists between calls (no cudaDeviceReset per call, :53); the arithmetic
quirks listed in the flags of ctr.h are off unless ref_quirks=True.
"""
import numpy as np

from . import _lib
from ._lib import ConeGeom, CurvedGeom, FanGeom, FdkErrGeom, ParGeom, default_context


def _col(a):
    return np.asarray(a, dtype=np.float64).reshape(-1)


def _flags(ref_quirks):
    return _lib.REF_QUIRKS if ref_quirks else 0


def ProjectionCone_3D(pic, Nt, Ns, Nz, Size, BetaScanRange, Pdomain, Xigamadomain, Distance, ctx=None):
    """R = ProjectionCone_3D(vol, Nt, Ns, Nz, Size, beta, Pdomain, Xigamadomain, Distance)
    (GPU_accelerated_Operator/ProjectionCone_3D/ProjectionCone_3D.cpp:3-59)."""
    ctx = ctx or default_context()
    g = ConeGeom(Nt, Ns, Nz, _col(Size)[:3], float(Distance), BetaScanRange, Pdomain, Xigamadomain)
    return ctx.cone_fp(g, pic).astype(np.float64)


def Backprojection(R, Xigamadomain, Pdomain, BetaScanRange, Distance, Size, Nt, Ns, Nz, ctx=None):
    """V = Backprojection(R, Xigamadomain, Pdomain, beta, Distance, Size, Nt, Ns, Nz)
    (GPU_accelerated_Operator/BackprojectionCone_3D/Backprojection3D.cpp:3-60)."""
    ctx = ctx or default_context()
    g = ConeGeom(Nt, Ns, Nz, _col(Size)[:3], float(Distance), BetaScanRange, Pdomain, Xigamadomain)
    return ctx.cone_bp_ref(g, R).astype(np.float64)


BackprojectionCone_3D = Backprojection
Backprojection3D = Backprojection


def FDK(R, Xigamadomain, Pdomain, BetaScanRange, Distance, Size, Nt, Ns, Nz, ref_quirks=False,
        filt=_lib.FILTER_SL_HALF, ctx=None):
    """V = FDK(R, Xigamadomain, Pdomain, beta, Distance, Size, Nt, Ns, Nz)
    (Analytic_Algorithm/FBP_type_analytic_algorithm_GPU_cpp_sourcecode/FDK_GPU/FDK.cpp:3-61)."""
    ctx = ctx or default_context()
    g = ConeGeom(Nt, Ns, Nz, _col(Size)[:3], float(Distance), BetaScanRange, Pdomain, Xigamadomain)
    return ctx.fdk(g, R, filt=filt, flags=_flags(ref_quirks)).astype(np.float64)


def ProjectionFan_2D(picvector, Nt, Ns, Size, BetaScanRange, Pdomain, Distance, ctx=None):
    """R = ProjectionFan_2D(img, Nt, Ns, Size, beta, Pdomain, Distance)
    (GPU_accelerated_Operator/ProjectionFan_2D/ProjectionFan_2D.cpp:3-54)."""
    ctx = ctx or default_context()
    g = FanGeom(Nt, Ns, _col(Size)[:2], float(Distance), BetaScanRange, Pdomain)
    return ctx.fan_fp(g, picvector).astype(np.float64)


def BackprojectionFan2D(R, BetaScanRange, Pdomain, Size, Nt, Ns, Rscan, ref_quirks=False, ctx=None):
    """img = BackprojectionFan2D(R, beta, Pdomain, Size, Nt, Ns, Rscan)
    (GPU_accelerated_Operator/BackprojectionFan_2D/BackprojectionFan2D.cpp:3-66).
    beta and Pdomain are single in the reference; they are rounded to float32 here too."""
    ctx = ctx or default_context()
    g = FanGeom(Nt, Ns, _col(Size)[:2], float(Rscan), _col(BetaScanRange).astype(np.float32),
                _col(Pdomain).astype(np.float32))
    return ctx.fan_bp_ref(g, R, flags=_flags(ref_quirks)).astype(np.float64)


BackprojectionFan_2D = BackprojectionFan2D


def FBPfan(R, BetaScanRange, Pdomain, Size, Nt, Ns, Rscan, filt=_lib.FILTER_SL_HALF, ctx=None):
    """img = FBPfan(R, beta, Pdomain, Size, Nt, Ns, Rscan)
    (Analytic_Algorithm/FBP_type_analytic_algorithm_GPU_cpp_sourcecode/FBP_fan_GPU/FBPfan.cpp:3-69)."""
    ctx = ctx or default_context()
    g = FanGeom(Nt, Ns, _col(Size)[:2], float(Rscan), _col(BetaScanRange).astype(np.float32),
                _col(Pdomain).astype(np.float32))
    return ctx.fan_fbp(g, R, detector=_lib.FAN_FLAT, filt=filt).astype(np.float64)


def ProjectionParallel_2D(picvector, height, width, Size, thetaRange, t_range, ctx=None):
    """R = ProjectionParallel_2D(img, height, width, Size, theta, t)
    (GPU_accelerated_Operator/ProjectionParallel_2D/ProjectionParallel_2D.cpp:3-50).
    The image is [height][width] with width fastest; Size(1) is the vertical
    extent (ProjectionParallel_2D.cpp:9-10), so the t axis is the width."""
    ctx = ctx or default_context()
    sz = _col(Size)
    g = ParGeom(int(width), int(height), (float(sz[1]), float(sz[0])), thetaRange, t_range)
    return ctx.par_fp(g, picvector).astype(np.float64)


def FBPparallel(R, ThetaScanRange, Tdomain, Size, t_length, s_length, filt=_lib.FILTER_SL_HALF, ctx=None):
    """img = FBPparallel(R, theta, Tdomain, Size, t_length, s_length)
    (Analytic_Algorithm/FBP_type_analytic_algorithm_GPU_cpp_sourcecode/FBP_parallel_GPU/FBPparallel.cpp:3-64).
    theta and Tdomain are single in the reference; they are rounded to float32 here too."""
    ctx = ctx or default_context()
    g = ParGeom(int(t_length), int(s_length), _col(Size)[:2], _col(ThetaScanRange).astype(np.float32),
                _col(Tdomain).astype(np.float32))
    return ctx.par_fbp(g, R, filt=filt).astype(np.float64)


def ImplementUI(R, Xigamadomain, Gamadomain, BetaScanRange, betaStartAngle, MaxGama, Distance, Distance_s2d, Size,
                z_length, ref_quirks=False, ctx=None):
    """Display = ImplementUI(R, Xigamadomain, Gamadomain, BetaScanRange, StartAngle, GamaMax, Distance,
    Distance_s2d, Size, z_length) (My_toolbox/Implement_UnitedImaging_Data/ImplementUI.cpp:3-57):
    curved-detector short-scan FDK.  Size(1:2) are the in-plane pixel counts at
    1 mm, Size(3) the z extent (mm); all three are truncated to integers as the
    reference does (ImplementUI.cpp:20-22)."""
    ctx = ctx or default_context()
    sz = [int(v) for v in _col(Size)[:3]]
    g = CurvedGeom(sz[0], sz[1], int(z_length), (float(sz[0]), float(sz[1]), float(sz[2])), float(Distance),
                   float(Distance_s2d), BetaScanRange, Gamadomain, Xigamadomain, float(betaStartAngle),
                   float(MaxGama))
    return ctx.curved_fdk(g, R, flags=_flags(ref_quirks)).astype(np.float64)


def IterativeFDKerrCompute(PriorImage, t_length, s_length, z_length, Size, BetaScanRange, Udomain, Pdomain, Distance,
                           G, t, ctx=None):
    """ErrorSlice = IterativeFDKerrCompute(PriorImage, t_length, s_length, z_length, Size, BetaScanRange,
    Udomain, Pdomain, Distance, G, t) (My_novel_algorithm/IterativeFDK/IterativeFDKerrCompute.cpp:3-56);
    t is 1-based as in MATLAB."""
    ctx = ctx or default_context()
    a = FdkErrGeom(t_length, s_length, z_length, _col(Size)[:3], float(Distance), _col(BetaScanRange),
                   _col(Udomain), _col(Pdomain), _col(G), int(t) - 1)
    return ctx.fdk_error_slice(a, PriorImage).astype(np.float64)


def SIRT3D(R, Xigamadomain, Pdomain, BetaScanRange, Distance, Size, Nt, Ns, Nz, nIter, Lamda,
           x0=None, minLim=0.0, maxLim=np.inf, ctx=None):
    """[x, res] = SIRT3D(R, Xigamadomain, Pdomain, beta, Distance, Size, Nt, Ns, Nz, nIter, lambda[, x0, min, max]).
    No reference MEX exists: the loop is Iterative_algorithm/SIRT3D.m:67-92 with
    beta in radians.  Returns the volume and the residual history
    ||b - A x_k|| / ||b|| for k = 0..nIter."""
    ctx = ctx or default_context()
    g = ConeGeom(Nt, Ns, Nz, _col(Size)[:3], float(Distance), BetaScanRange, Pdomain, Xigamadomain)
    x, r2, _ = ctx.sirt3d(g, R, int(nIter), float(Lamda), float(minLim), float(maxLim),
                          x0=None if x0 is None else np.asarray(x0, np.float32))
    return x.astype(np.float64), r2
