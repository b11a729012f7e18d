"""CPU tests of the host package and the C ABI library (no GPU compute)."""

import ctypes
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "sgx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sgx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(sg):
    from sinograms_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = _header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(_lib.EXPORTS) == syms
    assert L.sgx_version() >= 100


def test_library_error_paths_without_gpu(sg):
    from sinograms_amd import _lib
    L = _lib.load()
    L.sgx_last_error.restype = ctypes.c_char_p
    assert L.sgx_last_error(None) == b"null context"
    # a null context is rejected before any device call
    assert L.sgx_back_par(None, None, 4, 4, 1, None, None, 1.0, 0.0, None, 4, None, 4, None,
                          None, 0) == _lib.SGX_ERR_ARG
    out = ctypes.c_void_p()
    assert L.sgx_ctx_create(-1, ctypes.byref(out)) == _lib.SGX_ERR_DEVICE


# --- geometry -------------------------------------------------------------------

def test_geometry_defaults_and_short_scans(sg):
    rg = sg.SinoPar()
    assert rg.dims() == (128, 200)
    fa = sg.SinoFanArc()
    assert (fa.nb, fa.na, fa.dsd, fa.dod, fa.orbit) == (128, 200, 512.0, 128.0, 360.0)
    sh = sg.SinoFanArc.short()
    assert sh.na == 70                                  # d-angle.jl doctest geometry
    c1 = sg.SinoFanArc(nb=888, d=1.0239, na=984, dsd=949.0, dod=408.0)
    assert c1.dso == 541.0
    assert math.degrees(c1.gamma_max()) == pytest.approx(27.42, abs=0.01)
    c1s = sg.SinoFanArc.short(nb=888, d=1.0239, na=984, dsd=949.0, dod=408.0)
    assert c1s.na == 642
    ct = sg.CtFanFlat(ns=512, nt=384, ds=1.2, dt=1.2, na=360, dsd=1200.0, dod=400.0)
    assert ct.dims() == (512, 384, 360)
    assert math.degrees(ct.gamma_max()) == pytest.approx(14.33, abs=0.01)
    assert ct.rfov() > 181.0
    assert sg.CtFanArc().dt == 1.0


def test_circle_mask_count_config_fan(sg):
    rg = sg.SinoFanArc(nb=888, d=1.0239, na=984, dsd=949.0, dod=408.0)
    ig = sg.ImageGeom((512, 512), (0.98, 0.98))
    assert rg.rfov() == pytest.approx(249.1, abs=0.05)
    assert int(ig.circle(rg.rfov()).sum()) == 202968


# --- plan-level host computations vs the oracle ------------------------------------

def test_filter_response_host_vs_oracle(sg, orc):
    for rg in (sg.SinoPar(nb=367, d=1.0), sg.SinoFanArc(nb=888, d=1.0239, dsd=949.0, dod=408.0),
               sg.CtFanFlat(ns=512, nt=8, ds=1.2, dsd=1200.0, dod=400.0)):
        for win in (sg.Window(), sg.Window(sg.Hann()), sg.Window(sg.Hamming(), 0.8)):
            Hk = sg.fbp_filter(rg, window=win)
            npad = len(Hk)
            ds = rg.d if hasattr(rg, "nb") else rg.ds
            if hasattr(rg, "is_arc") and rg.is_arc:
                h, _ = orc.ramp_arc(npad, ds, rg.dsd)
            else:
                h, _ = orc.ramp_flat(npad, ds)
            shape = {"NoWindow": "none", "Hann": "hann", "Hamming": "hamming"}[
                type(win.shape).__name__]
            ref = orc.fbp_filter(h.astype(np.float32), npad, ds,
                                 window=orc.fbp_window(shape, npad, win.cutoff))
            assert np.allclose(Hk, ref, rtol=2e-6, atol=1e-6 * np.abs(ref).max())


def test_window_and_angle_weight_doctests(sg):
    assert np.allclose(sg.fbp_window(sg.Window(sg.Hamming()), 4), [1.0, 0.54, 0.0, 0.54])
    assert sg.fbp_window(sg.Window(sg.WindowVect(np.arange(1, 9))), 8)[0] == 5
    assert sg._angle_weights(sg.SinoPar().ar(f32=True)) == np.float32(0.015707962)
    assert sg._angle_weights(sg.SinoFanArc.short().ar(f32=True)) == pytest.approx(
        np.float32(0.04842342), rel=2e-7)


def test_parker_par_host(sg, orc):
    rg = sg.SinoPar(orbit=270.0)
    pw = sg.parker_weight(rg)
    assert pw.shape == (1, rg.na) and pw.dtype == np.float32 and pw[0, 0] == 0
    assert np.allclose(pw, orc.parker_par(270.0, rg.angles()), atol=1e-7)
    assert sg.parker_weight(sg.SinoPar()).shape == (1, 1)
    with pytest.raises(ValueError):
        sg.parker_weight(sg.SinoPar(orbit=361.0))


def test_sino_weight_and_cyl_weight_host(sg, orc):
    rg = sg.SinoFanFlat(nb=64, d=2.0)
    assert np.allclose(sg.fbp_sino_weight(rg), orc.fan_sino_weight(rg.gamma(), rg.dso, rg.dsd))
    ct = sg.CtFanArc(ns=16, nt=6, ds=1.0, dt=1.5)
    assert np.allclose(sg.fdk_weight_cyl(ct), orc.fdk_weight_cyl(ct.s(), ct.t(), ct.dsd,
                                                                   ct.dso, True), rtol=1e-6)


def test_host_error_paths_match_reference(sg):
    rg = sg.SinoPar(nb=12, na=10)
    ig = sg.ImageGeom((8, 8))
    with pytest.raises(sg.SizeMismatch, match="sino size"):
        sg.fbp_back(rg, ig, np.zeros((11, 10), np.float32))        # back-par.jl:39
    with pytest.raises(ValueError, match="nb=12 \\+ extra=5 > npad=16"):
        sg.fbp_sino_filter(np.zeros((12, 3), np.float32), np.zeros(16, np.float32), extra=5)
    with pytest.raises(ValueError):
        sg.ramp_flat(7, 1.0)                                       # ramp.jl:68


def test_phantom_disk_sinogram_closed_form(sg):
    rg = sg.SinoPar(nb=1024, d=0.01, na=16)
    ell = np.array([[0.5, -0.25, 2.0, 2.0, 0.3, 1.5]])
    p = sg.phantom.sino_ellipses(ell, rg).numpy()
    # a disk's projection integrates to rho * area for every view
    assert np.allclose(p.sum(axis=0) * rg.d, 1.5 * math.pi * 4, rtol=2e-3)


def test_phantom_ellipsoid_cone_integral(sg):
    # the projection of an ellipsoid through its centre along a ray equals the chord
    rg = sg.CtFanFlat(ns=9, nt=7, ds=1.0, dt=1.0, na=1, dsd=100.0, dod=40.0)
    ell = np.array([[0.0, 0.0, 0.0, 5.0, 5.0, 5.0, 2.0]])
    p = sg.phantom.proj_ellipsoids(ell, rg).numpy()
    assert p[4, 3, 0] == pytest.approx(2.0 * 10.0)
