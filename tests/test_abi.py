"""The C ABI without a GPU: the library loads, exports every symbol that
include/ctr/ctr.h declares, and rejects invalid arguments before any
device work.  Runs on the CPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "ctr", "ctr.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ctr_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    import ctr
    L = ctr.lib()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(ctr.EXPORTS) == syms


def test_version_and_status_strings():
    import ctr
    assert "gfx950" in ctr.version()
    L = ctr.lib()
    assert L.ctr_status_string(0) == b"ok"
    assert L.ctr_status_string(1) == b"invalid argument"
    assert L.ctr_status_string(99) == b"unknown status"


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_no_device_is_reported():
    import ctr
    L = ctr.lib()
    h = ctypes.c_void_p()
    assert L.ctr_ctx_create(0, ctypes.byref(h)) == ctr._lib.ENODEV
    assert h.value is None
    assert b"no HIP device" in L.ctr_last_error(None)
    with pytest.raises(ctr.CtrError):
        ctr.Context(0)


def cone(**kw):
    import ctr
    d = dict(nt=16, ns=16, nz=8, size=(16.0, 16.0, 8.0), D=100.0,
             beta=np.linspace(0.1, 6.2, 10), p=0.5 * (np.arange(40) - 19.5), xi=0.5 * (np.arange(20) - 9.5))
    d.update(kw)
    return ctr.ConeGeom(d["nt"], d["ns"], d["nz"], d["size"], d["D"], d["beta"], d["p"], d["xi"])


def test_geometry_validation():
    import ctr
    EINVAL, ETOOBIG = ctr._lib.EINVAL, ctr._lib.ETOOBIG
    assert cone().check(True) == 0
    assert cone(nt=0).check() == EINVAL
    assert cone(size=(16.0, -1.0, 8.0)).check() == EINVAL
    assert cone(D=10.0).check() == EINVAL                       # source orbit inside the volume
    assert cone(D=float("nan")).check() == EINVAL
    assert cone(p=np.array([0.0, 1.0, 1.0, 2.0])).check() == EINVAL   # not strictly monotonic
    nonuni = np.array([0.0, 0.5, 1.0, 1.6, 2.0])
    assert cone(p=nonuni).check(False) == 0                    # projection reads per-sample positions
    assert cone(p=nonuni).check(True) == EINVAL                # backprojection needs a uniform pitch
    assert b"non-uniform" in ctr.lib().ctr_last_error(None)
    assert cone(xi=np.array([0.0])).check(True) == EINVAL      # BP needs 2 rows
    assert cone(beta=np.array([0.1, np.inf])).check() == EINVAL
    assert cone(nt=2048, ns=2048, nz=1024).check() == ETOOBIG


def test_null_context_is_rejected_before_device_work():
    import ctr
    L = ctr.lib()
    g = cone()
    v = np.zeros(g.nvox, np.float32)
    pj = np.zeros(g.nray, np.float32)
    for fn, a, b in [(L.ctr_cone_fp, v, pj), (L.ctr_cone_bp_ref, pj, v), (L.ctr_cone_bp_adjoint, pj, v),
                     (L.ctr_fdk_bp, pj, v)]:
        assert fn(None, ctypes.byref(g.c), a.ctypes.data, b.ctypes.data, 0) == ctr._lib.EINVAL
    assert b"context is NULL" in L.ctr_last_error(None)
    # geometry errors are reported first
    bad = cone(D=1.0)
    assert L.ctr_cone_fp(None, ctypes.byref(bad.c), v.ctypes.data, pj.ctypes.data, 0) == ctr._lib.EINVAL
    assert b"intersects the volume" in L.ctr_last_error(None)


def test_fan_validation():
    import ctr
    g = ctr.FanGeom(32, 32, (32.0, 32.0), 100.0, np.linspace(0.1, 6.2, 20), 0.5 * (np.arange(64) - 31.5))
    assert g.check(True) == 0
    gam = 0.01 * (np.arange(64) - 31.5)
    eq = ctr.FanGeom(32, 32, (32.0, 32.0), 100.0, np.linspace(0.1, 6.2, 20), 100.0 * np.tan(gam))
    assert eq.check(False) == 0
    assert eq.check(True) == ctr._lib.EINVAL   # uniform in gamma, not in p
    assert ctr.FanGeom(32, 32, (32.0, 32.0), 20.0, [0.1], [0.0, 1.0]).check() == ctr._lib.EINVAL


def test_physical_mapping():
    """Iso-centre pitch = pitch * SOD / SDD; beta starts at dbeta (FDK_GPU.m:63-67);
    equiangular bins p = SOD tan(gamma)."""
    import ctr
    g = ctr.geometry.config_cone_256(360)
    assert np.allclose(np.diff(g.p), 0.5) and np.allclose(g.p, -g.p[::-1])
    assert g.p[0] == -127.75 and g.xi[-1] == 127.75
    assert np.isclose(g.beta[0], np.pi / 180, rtol=1e-6) and np.isclose(g.beta[-1], 2 * np.pi, rtol=1e-6)
    f = ctr.geometry.config_fan_2d()
    gam = np.arctan(f.p / 500.0)
    assert np.allclose(np.diff(gam), 1e-3) and len(f.p) == 512
    c5 = ctr.geometry.config_cone_512()
    assert np.isclose(c5.p[1] - c5.p[0], 0.8 * 700 / 1100, rtol=1e-4) and len(c5.p) == 1024 and len(c5.xi) == 768


def test_parallel_validation():
    import ctr
    EINVAL, ETOOBIG = ctr._lib.EINVAL, ctr._lib.ETOOBIG
    th, t = np.deg2rad(np.arange(1, 181)), 0.5 * (np.arange(90) - 44.5)
    assert ctr.ParGeom(32, 32, (32.0, 32.0), th, t).check(True) == 0
    assert ctr.ParGeom(0, 32, (32.0, 32.0), th, t).check() == EINVAL
    assert ctr.ParGeom(32, 32, (32.0, 32.0), [0.1, np.nan], t).check() == EINVAL
    assert ctr.ParGeom(32, 32, (32.0, 32.0), th, [0.0, 1.0, 1.0]).check() == EINVAL        # not monotonic
    t_nu = t.copy()
    t_nu[10] += 0.2
    assert ctr.ParGeom(32, 32, (32.0, 32.0), th, t_nu).check(False) == 0
    assert ctr.ParGeom(32, 32, (32.0, 32.0), th, t_nu).check(True) == EINVAL             # FBP needs a uniform pitch
    assert ctr.ParGeom(32, 32, (32.0, 32.0), np.zeros(70000), t).check() == ETOOBIG
    assert ctr.ParGeom(40000, 40000, (1.0, 1.0), th, t).check() == ETOOBIG


def test_curved_validation():
    import ctr
    EINVAL, ETOOBIG = ctr._lib.EINVAL, ctr._lib.ETOOBIG
    gam = 0.002 * (np.arange(101) - 50)
    xi = np.arange(8) - 3.5
    beta = 0.02 * np.arange(200)

    def geom(**kw):
        d = dict(nt=32, ns=32, nz=8, size=(32.0, 32.0, 8.0), D=100.0, Dsd=200.0, beta=beta, gamma=gam, xi=xi,
                 start=0.0, fan=0.2)
        d.update(kw)
        return ctr.CurvedGeom(d["nt"], d["ns"], d["nz"], d["size"], d["D"], d["Dsd"], d["beta"], d["gamma"],
                              d["xi"], d["start"], d["fan"])
    assert geom().check() == 0
    assert geom(D=10.0).check() == EINVAL                                  # source orbit inside the volume
    assert geom(fan=0.0).check() == EINVAL and geom(fan=4.0).check() == EINVAL
    assert geom(gamma=0.001 * (np.arange(1100) - 549.5)).check() == ETOOBIG   # FFT filter limit
    g2 = gam.copy()
    g2[3] += 1e-4
    assert geom(gamma=g2).check() == EINVAL                                 # non-uniform channels
    assert geom(gamma=np.array([0.0, 2.0])).check() == EINVAL              # outside (-pi/2, pi/2)
    assert geom(beta=np.array([0.0])).check() == EINVAL


def test_error_slice_and_helpers_reject_bad_arguments_without_a_context():
    import ctr
    from ctr.geometry import iterative_fdk_args
    L = ctr.lib()
    EINVAL = ctr._lib.EINVAL
    a = iterative_fdk_args(8, n_views=6, pint=4.0)
    vol = np.zeros(8 ** 3, np.float32)
    out = np.zeros(6 * 8 * 8, np.float32)
    bad_t = ctr.FdkErrGeom(8, 8, 8, a["size"], a["D"], a["beta"], a["u"], a["p"], a["g"], 8)
    assert L.ctr_fdk_error_slice(None, ctypes.byref(bad_t.c), vol.ctypes.data, out.ctypes.data, 0) == EINVAL
    assert b"slice t" in L.ctr_last_error(None)
    near = ctr.FdkErrGeom(8, 8, 8, a["size"], 20.0, a["beta"], a["u"], a["p"], a["g"], 3)
    assert L.ctr_fdk_error_slice(None, ctypes.byref(near.c), vol.ctypes.data, out.ctypes.data, 0) == EINVAL
    # tap indices are 16-bit on the device
    many = ctr.FdkErrGeom(8, 8, 8, a["size"], a["D"], a["beta"], a["u"], a["p"], np.zeros(32768), 3)
    assert L.ctr_fdk_error_slice(None, ctypes.byref(many.c), vol.ctypes.data, out.ctypes.data, 0) == ctr._lib.ETOOBIG
    assert b"32767 taps" in L.ctr_last_error(None)
    ok = ctr.FdkErrGeom(8, 8, 8, a["size"], a["D"], a["beta"], a["u"], a["p"], a["g"], 3)
    assert L.ctr_fdk_error_slice(None, ctypes.byref(ok.c), vol.ctypes.data, out.ctypes.data, 0) == EINVAL
    assert b"context is NULL" in L.ctr_last_error(None)
    ell = np.array([[1.0, 0.5, 0.5, -0.1, 0, 0, 0, 0]])
    assert L.ctr_phantom_ellipsoids(None, ell.ctypes.data, 1, 8, 8, 8, vol.ctypes.data, 0) == EINVAL
    assert b"semi-axis" in L.ctr_last_error(None)
    v = ctypes.c_double()
    assert L.ctr_rrmse(None, vol.ctypes.data, vol.ctypes.data, 0, ctypes.byref(v), 0) == EINVAL
