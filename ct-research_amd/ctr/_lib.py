"""ctypes binding of libctr.so (include/ctr/ctr.h).

The library must be built (ct-research_amd/build_lib.py); there is no
Python or CPU fallback for any operator.  Data arguments are numpy arrays
(host memory: the call stages them through the GPU and is synchronous) or
torch CUDA tensors (device memory: the call is ordered on the current
torch stream).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libctr.so")

OK, EINVAL, ENOMEM, EHIP, ETOOBIG, EFAULT_INJECTED, ENODEV = range(7)
DEVICE_PTRS, ACCUMULATE, REF_QUIRKS, AUTOTUNE = 0x1, 0x2, 0x4, 0x8
FILTER_SL_HALF, FILTER_RL_HALF = 0, 1
FAN_FLAT, FAN_EQUIANGULAR = 0, 1

# every symbol declared in include/ctr/ctr.h
EXPORTS = [
    "ctr_version", "ctr_status_string", "ctr_last_error", "ctr_ctx_create", "ctr_ctx_destroy",
    "ctr_ctx_set_stream", "ctr_ctx_stream", "ctr_ctx_synchronize", "ctr_device_alloc", "ctr_device_free",
    "ctr_copy_to_device", "ctr_copy_to_host", "ctr_cone_geom_check", "ctr_fan_geom_check",
    "ctr_cone_fp", "ctr_cone_bp_ref", "ctr_cone_bp_adjoint", "ctr_cone_filter", "ctr_fdk_bp", "ctr_fdk",
    "ctr_fan_fp", "ctr_fan_bp_ref", "ctr_fan_bp_adjoint", "ctr_fan_filter", "ctr_fan_fbp",
    "ctr_par_geom_check", "ctr_par_fp", "ctr_par_bp_adjoint", "ctr_par_filter", "ctr_par_fbp",
    "ctr_curved_geom_check", "ctr_curved_filter", "ctr_curved_bp", "ctr_curved_fdk",
    "ctr_cone_sysmat", "ctr_fan_sysmat", "ctr_par_sysmat", "ctr_fdk_error_slice",
    "ctr_phantom_ellipsoids", "ctr_rrmse", "ctr_local_error",
    "ctr_sirt3d", "ctr_sirt2d",
    "ctr_cnn_create", "ctr_cnn_destroy", "ctr_cnn_set_params", "ctr_cnn_get_params", "ctr_cnn_forward",
    "ctr_cnn_train_step", "ctr_sirt3d_step", "ctr_learned_sirt3d", "ctr_sirt3d_batched", "ctr_autotune_cone",
    "ctr_cone_bp_plan_stats", "ctr_cone_filter_noalloc", "ctr_cnn_forward_tiled", "ctr_sirt3d_emulated",
]

CNN_NUM_PARAMS = 57985
# (out, in) channels of the four 3x3x3 layers; weights then bias per layer in the flat vector
CNN_LAYERS = ((32, 2), (32, 32), (32, 32), (1, 32))


class CtrError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"ctr error {status}: {msg}")
        self.status = status


class ConeGeomC(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int), ("ns", ctypes.c_int), ("nz", ctypes.c_int),
                ("size", ctypes.c_double * 3), ("distance", ctypes.c_double),
                ("n_beta", ctypes.c_int), ("n_p", ctypes.c_int), ("n_xi", ctypes.c_int),
                ("beta", ctypes.c_void_p), ("p", ctypes.c_void_p), ("xi", ctypes.c_void_p)]


class FanGeomC(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int), ("ns", ctypes.c_int), ("size", ctypes.c_double * 2),
                ("distance", ctypes.c_double), ("n_beta", ctypes.c_int), ("n_p", ctypes.c_int),
                ("beta", ctypes.c_void_p), ("p", ctypes.c_void_p)]


class ParGeomC(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int), ("ns", ctypes.c_int), ("size", ctypes.c_double * 2),
                ("n_theta", ctypes.c_int), ("n_t", ctypes.c_int),
                ("theta", ctypes.c_void_p), ("t", ctypes.c_void_p)]


class CurvedGeomC(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int), ("ns", ctypes.c_int), ("nz", ctypes.c_int),
                ("size", ctypes.c_double * 3), ("distance", ctypes.c_double), ("distance_s2d", ctypes.c_double),
                ("n_beta", ctypes.c_int), ("n_gamma", ctypes.c_int), ("n_xi", ctypes.c_int),
                ("beta", ctypes.c_void_p), ("gamma", ctypes.c_void_p), ("xi", ctypes.c_void_p),
                ("beta_start", ctypes.c_double), ("fan_angle", ctypes.c_double)]


class FdkErrGeomC(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int), ("ns", ctypes.c_int), ("nz", ctypes.c_int), ("size", ctypes.c_double * 3),
                ("distance", ctypes.c_double), ("n_beta", ctypes.c_int), ("n_u", ctypes.c_int),
                ("n_p", ctypes.c_int), ("n_g", ctypes.c_int), ("beta", ctypes.c_void_p), ("u", ctypes.c_void_p),
                ("p", ctypes.c_void_p), ("g", ctypes.c_void_p), ("t", ctypes.c_int)]


class SirtOptsC(ctypes.Structure):
    _fields_ = [("n_iter", ctypes.c_int), ("lambda_", ctypes.c_float), ("min_val", ctypes.c_float),
                ("max_val", ctypes.c_float), ("use_x0", ctypes.c_int),
                ("residual_l2", ctypes.c_void_p), ("residual_l1", ctypes.c_void_p)]


class TuneResultC(ctypes.Structure):
    _fields_ = [("bp_pitch", ctypes.c_int), ("adj_cm", ctypes.c_int), ("fp_transpose", ctypes.c_int),
                ("bp_ms", ctypes.c_double * 3), ("adj_ms", ctypes.c_double * 3), ("fp_ms", ctypes.c_double * 2)]


class SirtProblemC(ctypes.Structure):
    _fields_ = [("geom", ConeGeomC), ("b", ctypes.c_void_p), ("x", ctypes.c_void_p),
                ("residual_l2", ctypes.c_void_p)]


class LearnedOptsC(ctypes.Structure):
    _fields_ = [("n_steps", ctypes.c_int), ("lambda_", ctypes.c_float), ("min_val", ctypes.c_float),
                ("max_val", ctypes.c_float), ("use_x0", ctypes.c_int),
                ("residual_l2", ctypes.c_void_p), ("accepted", ctypes.c_void_p), ("judge_vs_plain", ctypes.c_int),
                ("cnn_tile", ctypes.c_int)]


class EmulatedOptsC(ctypes.Structure):
    _fields_ = [("n_start", ctypes.c_int), ("n_target", ctypes.c_int), ("degree", ctypes.c_int),
                ("n_polish", ctypes.c_int), ("lambda_", ctypes.c_float), ("min_val", ctypes.c_float),
                ("max_val", ctypes.c_float), ("support_threshold", ctypes.c_float), ("cnn_tile", ctypes.c_int),
                ("support_degree", ctypes.c_int), ("residual_l2", ctypes.c_void_p), ("fell_back", ctypes.c_void_p)]


_lib = None


def lib():
    """Load libctr.so; raises if the HIP library has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libctr.so is not built ({LIB_PATH}); run ct-research_amd/build_lib.py")
        L = ctypes.CDLL(LIB_PATH)
        L.ctr_version.restype = ctypes.c_char_p
        L.ctr_status_string.restype = ctypes.c_char_p
        L.ctr_status_string.argtypes = [ctypes.c_int]
        L.ctr_last_error.restype = ctypes.c_char_p
        L.ctr_last_error.argtypes = [ctypes.c_void_p]
        L.ctr_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.ctr_ctx_destroy.argtypes = [ctypes.c_void_p]
        L.ctr_ctx_destroy.restype = None
        L.ctr_ctx_set_stream.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.ctr_ctx_stream.argtypes = [ctypes.c_void_p]
        L.ctr_ctx_stream.restype = ctypes.c_void_p
        L.ctr_ctx_synchronize.argtypes = [ctypes.c_void_p]
        L.ctr_cone_geom_check.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ctr_fan_geom_check.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ctr_par_geom_check.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ctr_curved_geom_check.argtypes = [ctypes.c_void_p]
        L.ctr_phantom_ellipsoids.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint]
        for name in ("ctr_rrmse", "ctr_local_error"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                         ctypes.POINTER(ctypes.c_double), ctypes.c_uint]
        L.ctr_fdk_error_slice.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_uint]
        for name in ("ctr_cone_sysmat", "ctr_fan_sysmat", "ctr_par_sysmat"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.c_uint]
        for name in ("ctr_cone_fp", "ctr_cone_bp_ref", "ctr_cone_bp_adjoint", "ctr_fdk_bp",
                     "ctr_fan_fp", "ctr_fan_bp_ref", "ctr_fan_bp_adjoint", "ctr_par_fp", "ctr_par_bp_adjoint",
                     "ctr_curved_filter", "ctr_curved_bp", "ctr_curved_fdk"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_uint]
        for name in ("ctr_cone_filter", "ctr_fdk", "ctr_par_filter", "ctr_par_fbp"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_uint]
        for name in ("ctr_fan_filter", "ctr_fan_fbp"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint]
        for name in ("ctr_sirt3d", "ctr_sirt2d"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_uint]
        L.ctr_cnn_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
        L.ctr_cnn_destroy.argtypes = [ctypes.c_void_p]
        L.ctr_cnn_destroy.restype = None
        L.ctr_cnn_set_params.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.ctr_cnn_get_params.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.ctr_cnn_forward.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p, ctypes.c_uint]
        L.ctr_cnn_forward_tiled.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_uint]
        L.ctr_cone_bp_plan_stats.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.ctr_cone_filter_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_uint]
        L.ctr_cnn_train_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_uint]
        L.ctr_autotune_cone.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.ctr_sirt3d_batched.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_uint]
        L.ctr_sirt3d_step.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_float, ctypes.c_uint]
        L.ctr_sirt3d_emulated.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint]
        L.ctr_learned_sirt3d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint]
        _lib = L
    return _lib


def version():
    return lib().ctr_version().decode()


def _f32(a):
    return np.ascontiguousarray(np.asarray(a).reshape(-1), dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a).reshape(-1), dtype=np.float64)


class ConeGeom:
    """Cone-beam geometry in the reference's argument conventions
    (ProjectionCone_3D.cpp:5-20): volume dims, Size (mm), source distance,
    view angles (rad), detector column / row positions at the isocentre."""

    def __init__(self, nt, ns, nz, size, distance, beta, p, xi):
        self.nt, self.ns, self.nz = int(nt), int(ns), int(nz)
        self.size = tuple(float(s) for s in size)
        self.distance = float(distance)
        self.beta, self.p, self.xi = _f32(beta), _f32(p), _f32(xi)
        self.c = ConeGeomC(self.nt, self.ns, self.nz, (ctypes.c_double * 3)(*self.size), self.distance,
                           len(self.beta), len(self.p), len(self.xi),
                           self.beta.ctypes.data, self.p.ctypes.data, self.xi.ctypes.data)

    @property
    def nvox(self):
        return self.nt * self.ns * self.nz

    @property
    def nray(self):
        return len(self.beta) * len(self.p) * len(self.xi)

    @property
    def vol_shape(self):
        return (self.nz, self.ns, self.nt)

    @property
    def proj_shape(self):
        return (len(self.beta), len(self.xi), len(self.p))

    def check(self, need_uniform=False):
        return lib().ctr_cone_geom_check(ctypes.byref(self.c), int(need_uniform))


class FanGeom:
    """Fan-beam geometry (ProjectionFan_2D.cpp:5-16)."""

    def __init__(self, nt, ns, size, distance, beta, p):
        self.nt, self.ns = int(nt), int(ns)
        self.size = tuple(float(s) for s in size)
        self.distance = float(distance)
        self.beta, self.p = _f64(beta), _f64(p)
        self.c = FanGeomC(self.nt, self.ns, (ctypes.c_double * 2)(*self.size), self.distance,
                          len(self.beta), len(self.p), self.beta.ctypes.data, self.p.ctypes.data)

    @property
    def nvox(self):
        return self.nt * self.ns

    @property
    def nray(self):
        return len(self.beta) * len(self.p)

    def check(self, need_uniform=False):
        return lib().ctr_fan_geom_check(ctypes.byref(self.c), int(need_uniform))


class ParGeom:
    """Parallel-beam geometry: image nt x ns pixels (t = x fastest) over size
    (mm), ray (theta, t) = the line (x - Cx) cos(theta) + (y - Cy) sin(theta) = t
    (ProjectionParallel_2D.cpp:5-16, FBPparallel.cpp:7-14)."""

    def __init__(self, nt, ns, size, theta, t):
        self.nt, self.ns = int(nt), int(ns)
        self.size = tuple(float(s) for s in size)
        self.theta, self.t = _f64(theta), _f64(t)
        self.c = ParGeomC(self.nt, self.ns, (ctypes.c_double * 2)(*self.size), len(self.theta), len(self.t),
                          self.theta.ctypes.data, self.t.ctypes.data)

    @property
    def nvox(self):
        return self.nt * self.ns

    @property
    def nray(self):
        return len(self.theta) * len(self.t)

    def check(self, need_uniform=False):
        return lib().ctr_par_geom_check(ctypes.byref(self.c), int(need_uniform))


class CurvedGeom:
    """Curved (cylindrical, equiangular) cone-beam short scan: volume nt x ns x nz
    over size (mm), projections [beta][xi][gamma] with gamma the channel angles
    (rad) and xi the row positions (mm) on the cylinder of radius distance_s2d;
    Parker window from beta_start over fan_angle (ImplementUI.cpp:7-24)."""

    def __init__(self, nt, ns, nz, size, distance, distance_s2d, beta, gamma, xi, beta_start, fan_angle):
        self.nt, self.ns, self.nz = int(nt), int(ns), int(nz)
        self.size = tuple(float(v) for v in size)
        self.distance, self.distance_s2d = float(distance), float(distance_s2d)
        self.beta, self.gamma, self.xi = _f64(beta), _f64(gamma), _f64(xi)
        self.beta_start, self.fan_angle = float(beta_start), float(fan_angle)
        self.c = CurvedGeomC(self.nt, self.ns, self.nz, (ctypes.c_double * 3)(*self.size), self.distance,
                             self.distance_s2d, len(self.beta), len(self.gamma), len(self.xi), self.beta.ctypes.data,
                             self.gamma.ctypes.data, self.xi.ctypes.data, self.beta_start, self.fan_angle)

    @property
    def nvox(self):
        return self.nt * self.ns * self.nz

    @property
    def nray(self):
        return len(self.beta) * len(self.gamma) * len(self.xi)

    def check(self):
        return lib().ctr_curved_geom_check(ctypes.byref(self.c))


class FdkErrGeom:
    """Arguments of the iterative-FDK error slice (IterativeFDKerrCompute.cpp:5-28):
    volume nt x ns x nz over size, views beta and rows u (float32), detector
    samples p and ramp taps g (float64), slice t (0-based)."""

    def __init__(self, nt, ns, nz, size, distance, beta, u, p, g, t):
        self.nt, self.ns, self.nz = int(nt), int(ns), int(nz)
        self.size = tuple(float(v) for v in size)
        self.distance = float(distance)
        self.beta, self.u = _f32(beta).copy(), _f32(u).copy()
        self.p, self.g = _f64(p), _f64(g)
        self.t = int(t)
        self.c = FdkErrGeomC(self.nt, self.ns, self.nz, (ctypes.c_double * 3)(*self.size), self.distance,
                             len(self.beta), len(self.u), len(self.p), len(self.g), self.beta.ctypes.data,
                             self.u.ctypes.data, self.p.ctypes.data, self.g.ctypes.data, self.t)

    @property
    def nvox(self):
        return self.nt * self.ns * self.nz

    @property
    def nout(self):
        return len(self.beta) * self.nz * self.ns


def _is_device(x):
    return hasattr(x, "data_ptr") and getattr(x, "is_cuda", False)


class Context:
    """A libctr context on one HIP device (stream, workspaces, cached tables)."""

    def __init__(self, device=0):
        L = lib()
        h = ctypes.c_void_p()
        st = L.ctr_ctx_create(int(device), ctypes.byref(h))
        if st != OK:
            raise CtrError(st, L.ctr_last_error(None).decode())
        self.h = h
        self._bound_stream = L.ctr_ctx_stream(h)   # the context's own stream

    def close(self):
        if self.h:
            lib().ctr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != OK:
            raise CtrError(st, lib().ctr_last_error(self.h).decode())

    def synchronize(self):
        self._check(lib().ctr_ctx_synchronize(self.h))

    def stream(self):
        return lib().ctr_ctx_stream(self.h)

    def set_stream(self, stream_ptr):
        self._check(lib().ctr_ctx_set_stream(self.h, stream_ptr))
        self._bound_stream = stream_ptr

    def _bind_torch_stream(self):
        import torch
        s = torch.cuda.current_stream().cuda_stream
        if s != self._bound_stream:
            self.set_stream(s)

    # -- argument plumbing ------------------------------------------------------
    def _io(self, inp, out_shape, out, accumulate=False):
        """Return (in_ptr, out_obj, out_ptr, flags)."""
        if _is_device(inp):
            import torch
            self._bind_torch_stream()
            if out is None:
                out = torch.zeros(out_shape, dtype=torch.float32, device=inp.device)
            assert out.dtype == torch.float32 and out.is_contiguous() and inp.is_contiguous()
            assert inp.dtype == torch.float32
            return inp.data_ptr(), out, out.data_ptr(), DEVICE_PTRS
        a = _f32(inp)
        if out is None:
            out = np.zeros(int(np.prod(out_shape)), dtype=np.float32)
        else:
            assert out.dtype == np.float32 and out.flags.c_contiguous
        return a, out, out.ctypes.data, 0

    def _run(self, fn, geom, inp, out_shape, out=None, flags=0, extra=()):
        a, out, optr, f = self._io(inp, out_shape, out)
        iptr = a if isinstance(a, int) else a.ctypes.data
        self._check(fn(self.h, ctypes.byref(geom.c), *extra, iptr, optr, f | flags))
        return out

    # -- cone -------------------------------------------------------------------
    def cone_fp(self, g, vol, out=None, flags=0):
        return self._run(lib().ctr_cone_fp, g, vol, g.nray, out, flags)

    def cone_bp_ref(self, g, proj, out=None, flags=0):
        return self._run(lib().ctr_cone_bp_ref, g, proj, g.nvox, out, flags)

    def cone_bp_adjoint(self, g, proj, out=None, flags=0):
        return self._run(lib().ctr_cone_bp_adjoint, g, proj, g.nvox, out, flags)

    def cone_filter(self, g, proj, filt=FILTER_SL_HALF, out=None, flags=0):
        return self._run(lib().ctr_cone_filter, g, proj, g.nray, out, flags, extra=(int(filt),))

    def fdk_bp(self, g, filtered, out=None, flags=0):
        return self._run(lib().ctr_fdk_bp, g, filtered, g.nvox, out, flags)

    def fdk(self, g, proj, filt=FILTER_SL_HALF, out=None, flags=0):
        return self._run(lib().ctr_fdk, g, proj, g.nvox, out, flags, extra=(int(filt),))

    # -- fan --------------------------------------------------------------------
    def fan_fp(self, g, img, out=None, flags=0):
        return self._run(lib().ctr_fan_fp, g, img, g.nray, out, flags)

    def fan_bp_ref(self, g, proj, out=None, flags=0):
        return self._run(lib().ctr_fan_bp_ref, g, proj, g.nvox, out, flags)

    def fan_bp_adjoint(self, g, proj, out=None, flags=0):
        return self._run(lib().ctr_fan_bp_adjoint, g, proj, g.nvox, out, flags)

    def fan_filter(self, g, proj, detector=FAN_FLAT, filt=FILTER_SL_HALF, out=None, flags=0):
        return self._run(lib().ctr_fan_filter, g, proj, g.nray, out, flags, extra=(int(detector), int(filt)))

    def fan_fbp(self, g, proj, detector=FAN_FLAT, filt=FILTER_SL_HALF, out=None, flags=0):
        return self._run(lib().ctr_fan_fbp, g, proj, g.nvox, out, flags, extra=(int(detector), int(filt)))

    # -- parallel ---------------------------------------------------------------
    def par_fp(self, g, img, out=None, flags=0):
        return self._run(lib().ctr_par_fp, g, img, g.nray, out, flags)

    def par_bp_adjoint(self, g, proj, out=None, flags=0):
        return self._run(lib().ctr_par_bp_adjoint, g, proj, g.nvox, out, flags)

    def par_filter(self, g, proj, filt=FILTER_SL_HALF, out=None, flags=0):
        return self._run(lib().ctr_par_filter, g, proj, g.nray, out, flags, extra=(int(filt),))

    def par_fbp(self, g, proj, filt=FILTER_SL_HALF, out=None, flags=0):
        return self._run(lib().ctr_par_fbp, g, proj, g.nvox, out, flags, extra=(int(filt),))

    # -- curved detector, short scan ----------------------------------------------
    def curved_filter(self, g, proj, out=None, flags=0):
        return self._run(lib().ctr_curved_filter, g, proj, g.nray, out, flags)

    def curved_bp(self, g, filtered, out=None, flags=0):
        return self._run(lib().ctr_curved_bp, g, filtered, g.nvox, out, flags)

    def curved_fdk(self, g, proj, out=None, flags=0):
        return self._run(lib().ctr_curved_fdk, g, proj, g.nvox, out, flags)

    # -- phantoms and metrics ------------------------------------------------------
    def phantom_ellipsoids(self, ell, nt, ns=None, nz=None, device=False):
        """Rasterize an [n][8] ellipsoid table (phantom3d.m) on the device; [nz][ns][nt]."""
        ns = nt if ns is None else ns
        nz = nt if nz is None else nz
        e = _f64(np.asarray(ell, np.float64).reshape(-1, 8))   # flat, 8 values per ellipsoid
        n_ell = e.size // 8
        if device:
            import torch
            self._bind_torch_stream()
            out = torch.empty(nt * ns * nz, dtype=torch.float32, device="cuda")
            self._check(lib().ctr_phantom_ellipsoids(self.h, e.ctypes.data, n_ell, nt, ns, nz, out.data_ptr(),
                                                     DEVICE_PTRS))
            return out.reshape(nz, ns, nt)
        out = np.empty(nt * ns * nz, np.float32)
        self._check(lib().ctr_phantom_ellipsoids(self.h, e.ctypes.data, n_ell, nt, ns, nz, out.ctypes.data, 0))
        return out.reshape(nz, ns, nt)

    def _metric(self, fn, x, ref):
        v = ctypes.c_double(0.0)
        if _is_device(x):
            self._bind_torch_stream()
            assert x.is_contiguous() and ref.is_contiguous() and x.numel() == ref.numel()
            self._check(fn(self.h, x.data_ptr(), ref.data_ptr(), x.numel(), ctypes.byref(v), DEVICE_PTRS))
        else:
            a, b = _f32(x).ravel(), _f32(ref).ravel()
            assert a.size == b.size
            self._check(fn(self.h, a.ctypes.data, b.ctypes.data, a.size, ctypes.byref(v), 0))
        return v.value

    def rrmse(self, x, ref):
        return self._metric(lib().ctr_rrmse, x, ref)

    def local_error(self, cur, prev):
        return self._metric(lib().ctr_local_error, cur, prev)

    # -- iterative FDK -------------------------------------------------------------
    def fdk_error_slice(self, a, vol, out=None, flags=0):
        return self._run(lib().ctr_fdk_error_slice, a, vol, a.nout, out, flags)

    # -- Siddon system matrix ----------------------------------------------------
    def _sysmat(self, fn, g, device):
        nnz = ctypes.c_int64(0)
        if device:
            import torch
            self._bind_torch_stream()
            row_ptr = torch.empty(g.nray + 1, dtype=torch.int64, device="cuda")
            self._check(fn(self.h, ctypes.byref(g.c), row_ptr.data_ptr(), None, None, ctypes.byref(nnz),
                           DEVICE_PTRS))
            col = torch.empty(max(nnz.value, 1), dtype=torch.int32, device="cuda")
            val = torch.empty(max(nnz.value, 1), dtype=torch.float32, device="cuda")
            self._check(fn(self.h, ctypes.byref(g.c), row_ptr.data_ptr(), col.data_ptr(), val.data_ptr(),
                           ctypes.byref(nnz), DEVICE_PTRS))
            return row_ptr, col[:nnz.value], val[:nnz.value]
        row_ptr = np.zeros(g.nray + 1, np.int64)
        self._check(fn(self.h, ctypes.byref(g.c), row_ptr.ctypes.data, None, None, ctypes.byref(nnz), 0))
        col = np.empty(max(nnz.value, 1), np.int32)
        val = np.empty(max(nnz.value, 1), np.float32)
        self._check(fn(self.h, ctypes.byref(g.c), row_ptr.ctypes.data, col.ctypes.data, val.ctypes.data,
                       ctypes.byref(nnz), 0))
        return row_ptr, col[:nnz.value], val[:nnz.value]

    def cone_sysmat(self, g, device=False):
        """CSR (row_ptr int64, col int32, val float32) of the Siddon system matrix."""
        return self._sysmat(lib().ctr_cone_sysmat, g, device)

    def fan_sysmat(self, g, device=False):
        return self._sysmat(lib().ctr_fan_sysmat, g, device)

    def par_sysmat(self, g, device=False):
        return self._sysmat(lib().ctr_par_sysmat, g, device)

    # -- SIRT -------------------------------------------------------------------
    def _sirt(self, fn, g, b, n_iter, lam, lo, hi, x0, residuals, flags):
        r2 = np.zeros(n_iter + 1) if residuals else None
        r1 = np.zeros(n_iter + 1) if residuals else None
        opts = SirtOptsC(int(n_iter), float(lam), float(lo), float(hi), int(x0 is not None),
                         r2.ctypes.data if residuals else None, r1.ctypes.data if residuals else None)
        if _is_device(b):
            import torch
            self._bind_torch_stream()
            x = x0 if x0 is not None else torch.zeros(g.nvox, dtype=torch.float32, device=b.device)
            self._check(fn(self.h, ctypes.byref(g.c), b.data_ptr(), x.data_ptr(), ctypes.byref(opts),
                           DEVICE_PTRS | flags))
        else:
            bb = _f32(b)
            x = _f32(x0).copy() if x0 is not None else np.zeros(g.nvox, dtype=np.float32)
            self._check(fn(self.h, ctypes.byref(g.c), bb.ctypes.data, x.ctypes.data, ctypes.byref(opts), flags))
        return x, r2, r1

    def sirt3d(self, g, b, n_iter, lam=1.0, lo=0.0, hi=np.inf, x0=None, residuals=True, flags=0):
        return self._sirt(lib().ctr_sirt3d, g, b, n_iter, lam, lo, hi, x0, residuals, flags)

    def sirt2d(self, g, b, n_iter, lam=1.0, lo=0.0, hi=np.inf, x0=None, residuals=True, flags=0):
        return self._sirt(lib().ctr_sirt2d, g, b, n_iter, lam, lo, hi, x0, residuals, flags)

    def autotune_cone(self, g):
        """Time the tile/layout variants on geometry g; the fastest are kept for later calls."""
        r = TuneResultC()
        self._check(lib().ctr_autotune_cone(self.h, ctypes.byref(g.c), ctypes.byref(r)))
        return {"bp_pitch": r.bp_pitch, "adj_cm": r.adj_cm, "fp_transpose": r.fp_transpose,
                "bp_ms": list(r.bp_ms), "adj_ms": list(r.adj_ms), "fp_ms": list(r.fp_ms)}

    def bp_plan_stats(self, g, adjoint=False):
        """Classes of the backprojection plan's (tile, view) pairs for geometry g."""
        c = (ctypes.c_longlong * 6)()
        self._check(lib().ctr_cone_bp_plan_stats(self.h, ctypes.byref(g.c), int(bool(adjoint)), c))
        return dict(zip(("pairs", "empty", "interior", "rows_in", "edge", "global"), (int(v) for v in c)))

    def sirt3d_batched(self, problems, n_iter, lam=1.0, lo=0.0, hi=np.inf, x0s=None, flags=0):
        """problems: [(ConeGeom, b)], b numpy or CUDA tensors (all alike).
        Returns ([x], [residual_l2]) with one graph-captured launch sequence per iteration."""
        n = len(problems)
        arr = (SirtProblemC * n)()
        res = [np.zeros(n_iter + 1) for _ in range(n)]
        xs, keep = [], []
        device = _is_device(problems[0][1])
        for i, (g, b) in enumerate(problems):
            if device:
                import torch
                x = x0s[i] if x0s is not None else torch.zeros(g.nvox, dtype=torch.float32, device=b.device)
                bp, xp = b.data_ptr(), x.data_ptr()
            else:
                bb = _f32(b)
                x = _f32(x0s[i]).copy() if x0s is not None else np.zeros(g.nvox, dtype=np.float32)
                keep.append(bb)
                bp, xp = bb.ctypes.data, x.ctypes.data
            xs.append(x)
            arr[i].geom = g.c
            arr[i].b = bp
            arr[i].x = xp
            arr[i].residual_l2 = res[i].ctypes.data
        opts = SirtOptsC(int(n_iter), float(lam), float(lo), float(hi), int(x0s is not None), None, None)
        if device:
            self._bind_torch_stream()
            flags |= DEVICE_PTRS
        self._check(lib().ctr_sirt3d_batched(self.h, arr, n, ctypes.byref(opts), flags))
        return xs, res

    # -- learned SIRT -------------------------------------------------------------
    def sirt3d_step(self, g, b, x, lam=1.0, out=None):
        """The plain SIRT step s = lam V A'(W (b - A x)) at x."""
        if _is_device(b):
            import torch
            self._bind_torch_stream()
            s = out if out is not None else torch.empty(g.nvox, dtype=torch.float32, device=b.device)
            self._check(lib().ctr_sirt3d_step(self.h, ctypes.byref(g.c), b.data_ptr(), x.data_ptr(), s.data_ptr(),
                                              float(lam), DEVICE_PTRS))
            return s
        bb, xx = _f32(b), _f32(x)
        s = np.zeros(g.nvox, dtype=np.float32)
        self._check(lib().ctr_sirt3d_step(self.h, ctypes.byref(g.c), bb.ctypes.data, xx.ctypes.data, s.ctypes.data,
                                          float(lam), 0))
        return s

    def sirt3d_emulated(self, g, b, net=None, n_start=5, n_target=210, degree=28, n_polish=2, lam=1.0, lo=0.0,
                        hi=np.inf, threshold=0.5, cnn_tile=0, support_degree=0, flags=0):
        """SIRT from zero: n_start plain iterations, the rest up to n_target
        emulated by a degree-`degree` Chebyshev polynomial on the support the
        network predicts (or the iterate's own), residual-judged, then
        n_polish plain iterations.  Returns (x, [res_start, res_emulated,
        res_end], fell_back)."""
        return _emulated(self, g, b, net, n_start, n_target, degree, n_polish, lam, lo, hi, threshold, cnn_tile,
                         support_degree, flags)

    def learned_sirt3d(self, g, b, net, n_steps, lam=1.0, lo=0.0, hi=np.inf, x0=None, judge_vs_plain=False, flags=0,
                       cnn_tile=0):
        """Residual-judged learned SIRT; returns (x, residual_l2[n_steps+1], accepted[n_steps]).
        cnn_tile > 0 runs the network in tiles of cnn_tile^3 voxels (8-voxel halos)."""
        r2 = np.zeros(n_steps + 1)
        acc = np.zeros(max(n_steps, 1), dtype=np.int32)
        opts = LearnedOptsC(int(n_steps), float(lam), float(lo), float(hi), int(x0 is not None),
                            r2.ctypes.data, acc.ctypes.data, int(bool(judge_vs_plain)), int(cnn_tile))
        fn = lib().ctr_learned_sirt3d
        if _is_device(b):
            import torch
            self._bind_torch_stream()
            x = x0 if x0 is not None else torch.zeros(g.nvox, dtype=torch.float32, device=b.device)
            self._check(fn(self.h, ctypes.byref(g.c), b.data_ptr(), x.data_ptr(), net.h, ctypes.byref(opts),
                           DEVICE_PTRS | flags))
        else:
            bb = _f32(b)
            x = _f32(x0).copy() if x0 is not None else np.zeros(g.nvox, dtype=np.float32)
            self._check(fn(self.h, ctypes.byref(g.c), bb.ctypes.data, x.ctypes.data, net.h, ctypes.byref(opts),
                           flags))
        return x, r2, acc[:n_steps]


def _emulated(ctx, g, b, net, n_start, n_target, degree, n_polish, lam, lo, hi, threshold, cnn_tile, support_degree,
              flags):
    res = np.zeros(3)
    fb = np.zeros(1, dtype=np.int32)
    opts = EmulatedOptsC(int(n_start), int(n_target), int(degree), int(n_polish), float(lam), float(lo), float(hi),
                         float(threshold), int(cnn_tile), int(support_degree), res.ctypes.data, fb.ctypes.data)
    fn = lib().ctr_sirt3d_emulated
    nh = net.h if net is not None else None
    if _is_device(b):
        import torch
        ctx._bind_torch_stream()
        x = torch.empty(g.nvox, dtype=torch.float32, device=b.device)
        ctx._check(fn(ctx.h, ctypes.byref(g.c), b.data_ptr(), x.data_ptr(), nh, ctypes.byref(opts), DEVICE_PTRS | flags))
    else:
        bb = _f32(b)
        x = np.zeros(g.nvox, dtype=np.float32)
        ctx._check(fn(ctx.h, ctypes.byref(g.c), bb.ctypes.data, x.ctypes.data, nh, ctypes.byref(opts), flags))
    return x, res, int(fb[0])


def cnn_init_params(seed=0, last_scale=0.0):
    """He-initialised parameters (zero biases); the last layer is scaled by
    last_scale (0: the network starts as the plain SIRT step)."""
    rng = np.random.default_rng(seed)
    out = []
    for li, (co, ci) in enumerate(CNN_LAYERS):
        std = np.sqrt(2.0 / (ci * 27))
        w = rng.standard_normal((co, ci, 3, 3, 3)) * std
        if li == len(CNN_LAYERS) - 1:
            w *= last_scale
        out += [w.ravel(), np.zeros(co)]
    p = np.concatenate(out).astype(np.float32)
    assert p.size == CNN_NUM_PARAMS
    return p


def cnn_split_params(p):
    """Flat parameters -> [(W[co][ci][3][3][3], b[co]), ...] for the four layers."""
    out, o = [], 0
    for co, ci in CNN_LAYERS:
        nw = co * ci * 27
        out.append((np.asarray(p[o:o + nw]).reshape(co, ci, 3, 3, 3), np.asarray(p[o + nw:o + nw + co])))
        o += nw + co
    return out


class Cnn:
    """The learned-update network on a context (ctr_cnn)."""

    def __init__(self, ctx, params=None):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx._check(lib().ctr_cnn_create(ctx.h, ctypes.byref(h)))
        self.h = h
        if params is not None:
            self.set_params(params)

    def close(self):
        if self.h:
            lib().ctr_cnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self.ctx._check(lib().ctr_cnn_set_params(self.h, p.ctypes.data, p.size))

    def get_params(self):
        p = np.zeros(CNN_NUM_PARAMS, dtype=np.float32)
        self.ctx._check(lib().ctr_cnn_get_params(self.h, p.ctypes.data, p.size))
        return p

    def forward(self, shape, x, g, g_scale=1.0):
        """shape = (nx, ny, nz); x, g: [nz][ny][nx] numpy or CUDA tensors."""
        nx, ny, nz = shape
        n = nx * ny * nz
        if _is_device(x):
            import torch
            self.ctx._bind_torch_stream()
            out = torch.empty(n, dtype=torch.float32, device=x.device)
            self.ctx._check(lib().ctr_cnn_forward(self.h, nx, ny, nz, x.data_ptr(), g.data_ptr(), float(g_scale),
                                                  out.data_ptr(), DEVICE_PTRS))
            return out
        xx, gg = _f32(x), _f32(g)
        out = np.zeros(n, dtype=np.float32)
        self.ctx._check(lib().ctr_cnn_forward(self.h, nx, ny, nz, xx.ctypes.data, gg.ctypes.data, float(g_scale),
                                              out.ctypes.data, 0))
        return out

    def forward_tiled(self, shape, x, g, g_scale=1.0, tile=128, halo=8):
        """forward() computed in tiles of tile^3 voxels with halo-voxel halos (identical result)."""
        nx, ny, nz = shape
        n = nx * ny * nz
        fn = lib().ctr_cnn_forward_tiled
        if _is_device(x):
            import torch
            self.ctx._bind_torch_stream()
            out = torch.empty(n, dtype=torch.float32, device=x.device)
            self.ctx._check(fn(self.h, nx, ny, nz, x.data_ptr(), g.data_ptr(), ctypes.c_float(g_scale), out.data_ptr(),
                               int(tile), int(halo), DEVICE_PTRS))
            return out
        xx, gg = _f32(x), _f32(g)
        out = np.zeros(n, dtype=np.float32)
        self.ctx._check(fn(self.h, nx, ny, nz, xx.ctypes.data, gg.ctypes.data, ctypes.c_float(g_scale), out.ctypes.data,
                           int(tile), int(halo), 0))
        return out

    def train_step(self, shape, x, g, g_scale, target, border=0, lr=0.0, want_grad=False):
        """Loss (and gradient) of a batch of patches [n][nz][ny][nx]; one Adam step when lr > 0."""
        nx, ny, nz = shape
        gs = np.ascontiguousarray(np.atleast_1d(g_scale), dtype=np.float32)
        npatch = gs.size
        loss = ctypes.c_double()
        grad = np.zeros(CNN_NUM_PARAMS, dtype=np.float32) if want_grad else None
        gptr = grad.ctypes.data if want_grad else None
        if _is_device(x):
            self.ctx._bind_torch_stream()
            self.ctx._check(lib().ctr_cnn_train_step(self.h, npatch, nx, ny, nz, x.data_ptr(), g.data_ptr(),
                                                     gs.ctypes.data, target.data_ptr(), int(border), float(lr),
                                                     ctypes.byref(loss), gptr, DEVICE_PTRS))
        else:
            xx, gg, tt = _f32(x), _f32(g), _f32(target)
            self.ctx._check(lib().ctr_cnn_train_step(self.h, npatch, nx, ny, nz, xx.ctypes.data, gg.ctypes.data,
                                                     gs.ctypes.data, tt.ctypes.data, int(border), float(lr),
                                                     ctypes.byref(loss), gptr, 0))
        return (loss.value, grad) if want_grad else loss.value


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx
