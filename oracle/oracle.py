"""ctypes binding of the FP64 CPU oracle (oracle/ctr_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker (or the timed CPU
baseline), never as the product path.  The product library (libctr.so)
never loads this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


class _Cone(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int), ("ns", ctypes.c_int), ("nz", ctypes.c_int),
                ("size", ctypes.c_double * 3), ("D", ctypes.c_double),
                ("nb", ctypes.c_int), ("np", ctypes.c_int), ("nx", ctypes.c_int),
                ("beta", ctypes.c_void_p), ("p", ctypes.c_void_p), ("xi", ctypes.c_void_p)]


class _Fan(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int), ("ns", ctypes.c_int), ("size", ctypes.c_double * 2),
                ("D", ctypes.c_double), ("nb", ctypes.c_int), ("np", ctypes.c_int),
                ("beta", ctypes.c_void_p), ("p", ctypes.c_void_p)]


class _Par(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int), ("ns", ctypes.c_int), ("size", ctypes.c_double * 2),
                ("nb", ctypes.c_int), ("np", ctypes.c_int), ("theta", ctypes.c_void_p), ("t", ctypes.c_void_p)]


class _Curved(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int), ("ns", ctypes.c_int), ("nz", ctypes.c_int), ("size", ctypes.c_double * 3),
                ("D", ctypes.c_double), ("Dsd", ctypes.c_double),
                ("nb", ctypes.c_int), ("ng", ctypes.c_int), ("nx", ctypes.c_int),
                ("beta", ctypes.c_void_p), ("gamma", ctypes.c_void_p), ("xi", ctypes.c_void_p),
                ("start", ctypes.c_double), ("fan", ctypes.c_double)]


class _FdkErr(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int), ("ns", ctypes.c_int), ("nz", ctypes.c_int), ("size", ctypes.c_double * 3),
                ("D", ctypes.c_double), ("nb", ctypes.c_int), ("nu", ctypes.c_int), ("np", ctypes.c_int),
                ("ng", ctypes.c_int), ("t", ctypes.c_int), ("beta", ctypes.c_void_p), ("u", ctypes.c_void_p),
                ("p", ctypes.c_void_p), ("g", ctypes.c_void_p)]


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Cone:
    """Cone geometry in the reference's argument conventions."""

    def __init__(self, nt, ns, nz, size, distance, beta, p, xi):
        self.beta, self.p, self.xi = _f32(beta), _f32(p), _f32(xi)
        self.s = _Cone(nt, ns, nz, (ctypes.c_double * 3)(*size), distance,
                       len(self.beta), len(self.p), len(self.xi),
                       _ptr(self.beta), _ptr(self.p), _ptr(self.xi))
        self.nvox = nt * ns * nz
        self.nray = len(self.beta) * len(self.p) * len(self.xi)


class Fan:
    def __init__(self, nt, ns, size, distance, beta, p):
        self.beta, self.p = _f64(beta), _f64(p)
        self.s = _Fan(nt, ns, (ctypes.c_double * 2)(*size), distance, len(self.beta), len(self.p),
                      _ptr(self.beta), _ptr(self.p))
        self.nvox = nt * ns
        self.nray = len(self.beta) * len(self.p)


class Par:
    """Parallel-beam geometry: image [ns][nt] (t fastest), size (t extent, s extent), theta (rad), t (mm)."""

    def __init__(self, nt, ns, size, theta, t):
        self.theta, self.t = _f64(theta), _f64(t)
        self.s = _Par(nt, ns, (ctypes.c_double * 2)(*size), len(self.theta), len(self.t),
                      _ptr(self.theta), _ptr(self.t))
        self.nvox = nt * ns
        self.nray = len(self.theta) * len(self.t)


class Curved:
    """Curved equiangular cone geometry: volume [nz][ns][nt], projections
    [beta][xi][gamma]; gamma channel angles (rad), xi row positions (mm) on the
    cylinder of radius Dsd; Parker window from start over the full fan angle."""

    def __init__(self, nt, ns, nz, size, D, Dsd, beta, gamma, xi, start, fan):
        self.beta, self.gamma, self.xi = _f64(beta), _f64(gamma), _f64(xi)
        self.s = _Curved(nt, ns, nz, (ctypes.c_double * 3)(*size), D, Dsd, len(self.beta), len(self.gamma),
                         len(self.xi), _ptr(self.beta), _ptr(self.gamma), _ptr(self.xi), start, fan)
        self.nvox = nt * ns * nz
        self.nray = len(self.beta) * len(self.gamma) * len(self.xi)


class FdkErr:
    """Arguments of the iterative-FDK error slice: volume [nz][ns][nt] over size,
    slice t (0-based), beta and u float32, p and the taps g float64."""

    def __init__(self, nt, ns, nz, size, D, beta, u, p, g, t):
        self.beta, self.u = _f32(beta), _f32(u)
        self.p, self.g = _f64(p), _f64(g)
        self.s = _FdkErr(nt, ns, nz, (ctypes.c_double * 3)(*size), D, len(self.beta), len(self.u), len(self.p),
                         len(self.g), t, _ptr(self.beta), _ptr(self.u), _ptr(self.p), _ptr(self.g))
        self.nvox = nt * ns * nz
        self.nout = len(self.beta) * nz * ns


def _call(name, g, *args):
    fn = getattr(lib(), name)
    fn.restype = None
    fn(ctypes.byref(g.s), *args)


def cone_fp_exact(g, vol):
    out = np.zeros(g.nray)
    v = _f32(vol)
    _call("or_cone_fp_exact", g, _ptr(v), _ptr(out))
    return out


def cone_fp_ref(g, vol):
    out = np.zeros(g.nray)
    v = _f32(vol)
    _call("or_cone_fp_ref", g, _ptr(v), _ptr(out))
    return out


def cone_rowsum(g):
    out = np.zeros(g.nray)
    _call("or_cone_rowsum", g, _ptr(out))
    return out


def cone_adjoint(g, proj):
    out = np.zeros(g.nvox)
    y = _f32(proj)
    _call("or_cone_adjoint", g, _ptr(y), _ptr(out))
    return out


def cone_ray_brute(g, vol, b, ip, ix):
    fn = lib().or_cone_ray_brute
    fn.restype = ctypes.c_double
    v = _f32(vol)
    return fn(ctypes.byref(g.s), _ptr(v), b, ip, ix)


def cone_bp_ref(g, proj, quirks=False):
    out = np.zeros(g.nvox)
    y = _f32(proj)
    _call("or_cone_bp_ref", g, _ptr(y), _ptr(out), ctypes.c_int(int(quirks)))
    return out


def cone_filter(g, proj, filt=0):
    out = np.zeros(g.nray)
    y = _f32(proj)
    _call("or_cone_filter", g, ctypes.c_int(filt), _ptr(y), _ptr(out))
    return out


def fdk_bp(g, filtered, quirks=False):
    out = np.zeros(g.nvox)
    y = _f32(filtered)
    _call("or_fdk_bp", g, _ptr(y), _ptr(out), ctypes.c_int(int(quirks)))
    return out


def filter_taps(filt, equiangular, d, n):
    out = np.zeros(n)
    fn = lib().or_filter_taps
    fn.restype = None
    fn(ctypes.c_int(filt), ctypes.c_int(int(equiangular)), ctypes.c_double(d), ctypes.c_int(n), _ptr(out))
    return out


def fan_fp_exact(g, img):
    out = np.zeros(g.nray)
    v = _f32(img)
    _call("or_fan_fp_exact", g, _ptr(v), _ptr(out))
    return out


def fan_fp_ref(g, img):
    out = np.zeros(g.nray)
    v = _f32(img)
    _call("or_fan_fp_ref", g, _ptr(v), _ptr(out))
    return out


def fan_rowsum(g):
    out = np.zeros(g.nray)
    _call("or_fan_rowsum", g, _ptr(out))
    return out


def fan_adjoint(g, proj):
    out = np.zeros(g.nvox)
    y = _f32(proj)
    _call("or_fan_adjoint", g, _ptr(y), _ptr(out))
    return out


def fan_bp_ref(g, proj, quirks=False):
    out = np.zeros(g.nvox)
    y = _f32(proj)
    _call("or_fan_bp_ref", g, _ptr(y), _ptr(out), ctypes.c_int(int(quirks)))
    return out


def fan_filter(g, proj, equiangular=False, filt=0):
    out = np.zeros(g.nray)
    y = _f32(proj)
    _call("or_fan_filter", g, ctypes.c_int(int(equiangular)), ctypes.c_int(filt), _ptr(y), _ptr(out))
    return out


def fan_fbp_bp(g, filtered, equiangular=False):
    out = np.zeros(g.nvox)
    y = _f32(filtered)
    _call("or_fan_fbp_bp", g, ctypes.c_int(int(equiangular)), _ptr(y), _ptr(out))
    return out


def par_fp_exact(g, img):
    out = np.zeros(g.nray)
    v = _f32(img)
    _call("or_par_fp_exact", g, _ptr(v), _ptr(out))
    return out


def par_adjoint(g, proj):
    out = np.zeros(g.nvox)
    y = _f32(proj)
    _call("or_par_adjoint", g, _ptr(y), _ptr(out))
    return out


def par_ray_brute(g, img, b, it):
    fn = getattr(lib(), "or_par_ray_brute")
    fn.restype = ctypes.c_double
    v = _f32(img)
    return fn(ctypes.byref(g.s), _ptr(v), ctypes.c_int(b), ctypes.c_int(it))


def par_filter(g, proj, filt=0):
    out = np.zeros(g.nray)
    y = _f32(proj)
    _call("or_par_filter", g, ctypes.c_int(filt), _ptr(y), _ptr(out))
    return out


def par_fbp_bp(g, filtered):
    out = np.zeros(g.nvox)
    y = _f32(filtered)
    _call("or_par_fbp_bp", g, _ptr(y), _ptr(out))
    return out


def _segments(name, g, ray, cap):
    fn = getattr(lib(), name)
    fn.restype = ctypes.c_int
    idx = np.zeros(cap, np.int64)
    ln = np.zeros(cap)
    m = fn(ctypes.byref(g.s), ctypes.c_long(ray), _ptr(idx), _ptr(ln))
    return idx[:m], ln[:m]


def cone_segments(g, ray):
    """Exact Siddon segments (voxel index, length) of one ray, in visiting order."""
    s = g.s
    return _segments("or_cone_segments", g, ray, s.nt + s.ns + s.nz + 8)


def fan_segments(g, ray):
    s = g.s
    return _segments("or_fan_segments", g, ray, s.nt + s.ns + 9)


def par_segments(g, ray):
    s = g.s
    return _segments("or_par_segments", g, ray, s.nt + s.ns + 9)


def fdk_error_slice(a, vol):
    out = np.zeros(a.nout)
    v = _f32(vol)
    _call("or_fdk_error_slice", a, _ptr(v), _ptr(out))
    return out


def parker(beta, start, gamma, fan):
    fn = lib().or_parker
    fn.restype = ctypes.c_double
    fn.argtypes = [ctypes.c_double] * 4
    return fn(beta, start, gamma, fan)


def curved_taps(g, x):
    out = np.zeros(len(g.gamma))
    _call("or_curved_taps", g, ctypes.c_int(x), _ptr(out))
    return out


def curved_filter(g, proj):
    out = np.zeros(g.nray)
    y = _f32(proj)
    _call("or_curved_filter", g, _ptr(y), _ptr(out))
    return out


def curved_bp(g, filtered, quirks=False):
    out = np.zeros(g.nvox)
    y = _f32(filtered)
    _call("or_curved_bp", g, _ptr(y), _ptr(out), ctypes.c_int(int(quirks)))
    return out


def sirt3d(g, b, n_iter, lam=1.0, lo=0.0, hi=np.inf, x0=None):
    x = np.zeros(g.nvox) if x0 is None else _f64(x0).copy()
    r2 = np.zeros(n_iter + 1)
    r1 = np.zeros(n_iter + 1)
    bb = _f32(b)
    _call("or_sirt3d", g, _ptr(bb), _ptr(x), ctypes.c_int(n_iter), ctypes.c_double(lam),
          ctypes.c_double(lo), ctypes.c_double(hi), _ptr(r2), _ptr(r1))
    return x, r2, r1


def sirt2d(g, b, n_iter, lam=1.0, lo=0.0, hi=np.inf, x0=None):
    x = np.zeros(g.nvox) if x0 is None else _f64(x0).copy()
    r2 = np.zeros(n_iter + 1)
    r1 = np.zeros(n_iter + 1)
    bb = _f32(b)
    _call("or_sirt2d", g, _ptr(bb), _ptr(x), ctypes.c_int(n_iter), ctypes.c_double(lam),
          ctypes.c_double(lo), ctypes.c_double(hi), _ptr(r2), _ptr(r1))
    return x, r2, r1


def set_interp_mode(mode):
    """0: the reference's first-pair detector grid; 1: the uniform grid through
    the end samples (what the HIP backprojectors interpolate on)."""
    fn = lib().or_set_interp_mode
    fn.restype = None
    fn(ctypes.c_int(int(mode)))


class interp_mode:
    """Context manager selecting the oracle's interpolation grid."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        set_interp_mode(self.mode)

    def __exit__(self, *exc):
        set_interp_mode(0)


def num_threads():
    fn = lib().or_num_threads
    fn.restype = ctypes.c_int
    return fn()


def set_num_threads(n):
    """Threads of the oracle's OpenMP loops (n > 0)."""
    fn = lib().or_set_num_threads
    fn.restype = None
    fn(ctypes.c_int(int(n)))
