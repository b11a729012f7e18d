"""CPU oracle for the Sinograms.jl FBP / FDK hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the HIP library and
the ``sinograms_amd`` host package) imports, links or calls this module.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker.

Two layers:

* numpy restatements of the plan-level formulas (geometry scalars, ramp,
  window, filter response, Parker, angle weights, fan / FDK weights), each
  citing the reference file:line it follows;
* ctypes bindings of ``oracle.cpp`` (built to ``oracle/build/liboracle.so``)
  for the hot loops: the three backprojectors, the row filter and the
  distance-driven projector pair, in Float64 and in reference-precision
  Float32 mode.

Arrays follow Julia's column-major convention.  A Julia array of size
(n1, n2, ...) is held here as a numpy array of that shape in Fortran order.

Parity pins for every function live in ``tests/test_oracle.py`` (closed forms,
doctest values of the reference, known-answer and invariant checks).  The
element-level Float32 rounding of FFTW's transforms and of Julia's libm is
not reproduced ("parity unpinned" at the ~1e-7 level); every tolerance used
against the oracle is far above that.
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile oracle.cpp with the Makefile next to this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_int = ctypes.c_int
_vp = ctypes.c_void_p


def _declare(L):
    for suf in ("f64", "f32"):
        getattr(L, f"orc_back_par_{suf}").argtypes = [
            _vp, _i64, _i64, _i64, _vp, _vp, _f64, _f64, _vp, _i64, _vp, _i64, _vp, _vp]
        getattr(L, f"orc_back_fan_{suf}").argtypes = [
            _vp, _i64, _i64, _i64, _vp, _vp, _f64, _f64, _f64, _int, _f64, _f64,
            _vp, _i64, _vp, _i64, _vp, _vp]
        getattr(L, f"orc_cbct_back_{suf}").argtypes = [
            _vp, _i64, _i64, _i64, _i64, _vp, _f64, _f64, _f64, _f64, _f64, _f64, _int,
            _vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp]
        getattr(L, f"orc_sino_filter_{suf}").argtypes = [
            _vp, _i64, _i64, _vp, _i64, _i64, _vp]
        getattr(L, f"orc_cbct_back_opts_{suf}").argtypes = [
            _vp, _i64, _i64, _i64, _i64, _vp, _f64, _f64, _f64, _f64, _f64, _f64, _int,
            _vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _int, _vp]
        getattr(L, f"orc_cbct_back_opts2_{suf}").argtypes = [
            _vp, _i64, _i64, _i64, _i64, _vp, _f64, _f64, _f64, _f64, _f64, _f64, _int,
            _vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _int, _f64, _vp]
        getattr(L, f"orc_cbct_back_rp_{suf}").argtypes = [
            _vp, _vp, _i64, _i64, _i64, _vp, _f64, _f64, _f64, _f64, _int,
            _vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp]
    L.orc_filter_kernel.argtypes = [_vp, _i64, _vp]
    L.orc_project_bdd.argtypes = [_vp, _i64, _i64, _vp, _i64, _f64, _f64, _f64, _f64, _vp]
    L.orc_backproject_bdd.argtypes = [_vp, _i64, _i64, _vp, _i64, _f64, _f64, _f64, _f64,
                                      _int, _vp]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _F(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


# ---------------------------------------------------------------------------
# geometry scalars
# ---------------------------------------------------------------------------

def angles_rad(na, orbit, orbit_start=0.0, f32=False):
    """``_ar(rg)`` (src/geom/common.jl:27-34, src/unit.jl:10).

    Julia forms the Float32 range in extended precision and rounds once; the
    Float32 mode here multiplies in double by Float32(deg2rad(1)) and rounds.
    """
    i = np.arange(na, dtype=np.float64)
    if f32:
        k = float(np.float32(math.pi / 180))
        deg = float(np.float32(orbit_start)) + i * (float(np.float32(orbit)) / na)
        return (deg * k).astype(np.float32)
    return (orbit_start + i * (orbit / na)) * (math.pi / 180)


def s_axis(n, d, offset=0.0):
    """``_s(rg)`` = d * ((0:n-1) .- ((n-1)/2 + offset))  (src/geom/sino-geom.jl:18-20)."""
    return d * (np.arange(n, dtype=np.float64) - ((n - 1) / 2 + offset))


def gamma(s, dsd, is_arc):
    """``_gamma`` (src/geom/common.jl:93-97)."""
    s = np.asarray(s, dtype=np.float64)
    return s / dsd if is_arc else np.arctan(s / dsd)


def orbit_short_deg(gamma_max):
    """``_orbit_short`` (src/geom/common.jl:40)."""
    return 180 + 2 * math.degrees(gamma_max)


# ---------------------------------------------------------------------------
# ramp, window, filter response  (src/fbp/ramp.jl, window.jl, filter.jl)
# ---------------------------------------------------------------------------

def ramp_flat(N, ds):
    """``ramp_flat`` (src/fbp/ramp.jl:233-238, 199-205); returns (h, n)."""
    if N % 2:
        raise ValueError("N must be even")
    n = np.arange(-(N // 2), N // 2)
    h = np.zeros(N)
    h[n == 0] = 0.25 / ds ** 2
    odd = n % 2 == 1
    h[odd] = -1.0 / (math.pi * n[odd] * ds) ** 2
    return h, n


def ramp_arc(N, ds, dsd):
    """``ramp_arc`` (src/fbp/ramp.jl:214-224, 191-197); returns (h, n)."""
    if N % 2:
        raise ValueError("N must be even")
    if N / 2 * ds / dsd > 0.9 * math.pi / 2:
        raise ValueError("physically impossible arc geometry")
    n = np.arange(-(N // 2), N // 2)
    h = np.zeros(N)
    h[n == 0] = 0.25 / ds ** 2
    odd = n % 2 == 1
    h[odd] = -1.0 / (math.pi * dsd * np.sin(n[odd] * ds / dsd)) ** 2
    return h, n


def fbp_window(shape, N, cutoff=1.0, vect=None):
    """``fbp_window`` (src/fbp/window.jl:60-91): window over n = -N/2..N/2-1,
    cast to Float32 and fftshift-ed (so entry 0 is n = 0)."""
    if shape == "vect":
        v = np.asarray(vect, dtype=np.float64)
        if len(v) != N:
            raise ValueError("N mismatch")
        return np.roll(v, -(N // 2))
    width = cutoff * N
    n = np.arange(N) - N // 2
    inside = np.abs(n) < width / 2
    if shape == "none":
        w = np.ones(N)
    elif shape == "boxcar":
        w = inside.astype(np.float64)
    elif shape == "hamming":
        w = (0.54 + 0.46 * np.cos(2 * math.pi * n / width)) * inside
    elif shape == "hann":
        w = 0.5 * (1 + np.cos(2 * math.pi * n / width)) * inside
    else:
        raise ValueError(shape)
    w = w.astype(np.float32).astype(np.float64)
    return np.roll(w, -(N // 2))   # fftshift for even N


def fbp_filter(h, npad, ds, window=None, decon1=True):
    """``fbp_filter`` (src/fbp/filter.jl:32-57) without an FFT library:
    Hk[j] = sum_n h[n] cos(2 pi j n / npad) is the real part of the DFT of the
    fftshift-ed even sequence h; then .* window, * ds, ./ fftshift(sinc^2)."""
    h = np.asarray(h, dtype=np.float64)
    n = np.arange(-(npad // 2), npad // 2)
    j = np.arange(npad)
    phase = np.outer(j, n) % npad             # exact integer phase reduction
    Hk = np.cos(2 * math.pi * phase / npad) @ h
    if window is not None:
        Hk = Hk * window
    Hk = ds * Hk
    if decon1:
        x = n / npad
        sinc2 = np.where(x == 0, 1.0, (np.sin(math.pi * x) / (math.pi * np.where(x == 0, 1, x))) ** 2)
        Hk = Hk / np.roll(sinc2, -(npad // 2))
    return Hk


def filter_kernel(Hk):
    """g = real(ifft(Hk)) by a direct inverse DFT in double (oracle.cpp)."""
    Hk = np.ascontiguousarray(Hk, dtype=np.float64)
    g = np.empty(len(Hk))
    lib().orc_filter_kernel(_p(Hk), len(Hk), _p(g))
    return g


# ---------------------------------------------------------------------------
# weights (src/fbp/parker.jl, d-angle.jl, fbp2/sino-weight.jl, fbp3/fdk.jl)
# ---------------------------------------------------------------------------

def angle_weight(ar):
    """``_angle_weights`` (src/fbp/d-angle.jl:209-220)."""
    ar = np.asarray(ar)
    da = np.diff(ar)
    if not np.allclose(da, da[0], rtol=math.sqrt(np.finfo(ar.dtype).eps), atol=0):
        raise ValueError("Unequal angles not implemented yet")
    last = ar[-1] - ar[-2]
    nxt = ar[-1] + last
    span = nxt - ar[0]
    n180 = max(int(math.floor(1.00001 * abs(float(span)) / math.pi)), 1)
    return abs(da[0] / ar.dtype.type(n180))


def parker_fan_short(gam, ar, gamma_max):
    """``parker_weight_fan_short`` (src/fbp/parker.jl:62-108), (nb, na).

    The three regions are evaluated in the reference's order; a sample at
    exactly beta == pi - 2 gamma gets 0 (parker.jl:101,104).
    """
    gam = np.asarray(gam, dtype=np.float64)
    ar = np.asarray(ar, dtype=np.float64)
    bet = np.maximum(ar - ar.min(), 0.0)
    gg, bb = np.meshgrid(gam, bet, indexing="ij")
    fun = lambda x: np.sin(math.pi / 2 * x) ** 2
    wt = np.zeros(gg.shape)
    r1 = bb < 2 * (gamma_max - gg)
    wt[r1] = fun(bb[r1] / (2 * (gamma_max - gg[r1])))
    r2 = (2 * (gamma_max - gg) <= bb) & (bb < math.pi - 2 * gg)
    wt[r2] = 1.0
    r3 = (math.pi - 2 * gg < bb) & (bb <= math.pi + 2 * gamma_max)
    wt[r3] = fun((math.pi + 2 * gamma_max - bb[r3]) / (2 * (gamma_max + gg[r3])))
    return wt


def parker_par(orbit, ad):
    """``parker_weight_par`` (src/fbp/parker.jl:26-55); ad in degrees."""
    ad = np.asarray(ad, dtype=np.float64)
    if (orbit // 180) * 180 == orbit:
        return np.ones((1, 1))
    if abs(orbit) < 180:
        return np.ones((1, 1))
    orbit = abs(orbit)
    if orbit > 360:
        raise ValueError("only 180 <= |orbit| <= 360 supported for Parker weighting")
    extra = orbit - 180
    wt = np.ones((1, len(ad)))
    a = np.abs(ad)
    ii = a < extra
    wt[0, ii] = np.sin(a[ii] / extra * math.pi / 2) ** 2
    ii = a >= 180
    wt[0, ii] = np.sin((orbit - a[ii]) / extra * math.pi / 2) ** 2
    return wt


def fan_sino_weight(gam, dso, dsd, source_offset=0.0):
    """``fbp_sino_weight`` (src/fbp2/sino-weight.jl:12-16), (nb,)."""
    gam = np.asarray(gam, dtype=np.float64)
    return np.abs(dso * np.cos(gam) - source_offset * np.sin(gam)) / dsd


def fdk_weight_cyl(s, t, dsd, dso, is_arc):
    """``fdk_weight_cyl_{arc,flat}`` (src/fbp3/fdk.jl:21-45), (ns, nt)."""
    S, T = np.meshgrid(np.asarray(s, np.float64), np.asarray(t, np.float64), indexing="ij")
    if is_arc:
        return (dso / dsd) * np.cos(S / dsd) / np.sqrt((T / dsd) ** 2 + 1)
    return dso / np.sqrt(S ** 2 + T ** 2 + dsd ** 2)


# ---------------------------------------------------------------------------
# hot loops (oracle.cpp)
# ---------------------------------------------------------------------------

def _suffix(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def back_par(sino, sinphi, cosphi, ds, offset, xc, yc, mask, ia_skip=1, dtype=np.float64):
    """``fbp_back_par!`` (src/fbp2/back-par.jl:206-253)."""
    sino = _F(sino, dtype)
    nb, na = sino.shape
    xc = _F(xc, np.float64); yc = _F(yc, np.float64)
    mask = _F(mask, np.uint8)
    sp = _F(sinphi, np.float64); cp = _F(cosphi, np.float64)
    img = np.zeros((len(xc), len(yc)), dtype=dtype, order="F")
    getattr(lib(), "orc_back_par_" + _suffix(dtype))(
        _p(sino), nb, na, ia_skip, _p(sp), _p(cp), ds, offset,
        _p(xc), len(xc), _p(yc), len(yc), _p(mask), _p(img))
    return img


def back_fan(sino, sinb, cosb, dsd, dso, source_offset, is_arc, ds, offset, xc, yc, mask,
             ia_skip=1, dtype=np.float64):
    """``fbp_back_fan!`` (src/fbp2/back-fan.jl:215-254)."""
    sino = _F(sino, dtype)
    nb, na = sino.shape
    xc = _F(xc, np.float64); yc = _F(yc, np.float64)
    mask = _F(mask, np.uint8)
    sb = _F(sinb, np.float64); cb = _F(cosb, np.float64)
    img = np.zeros((len(xc), len(yc)), dtype=dtype, order="F")
    getattr(lib(), "orc_back_fan_" + _suffix(dtype))(
        _p(sino), nb, na, ia_skip, _p(sb), _p(cb), dsd, dso, source_offset, int(bool(is_arc)),
        ds, offset, _p(xc), len(xc), _p(yc), len(yc), _p(mask), _p(img))
    return img


def cbct_back(proj, betas, dsd, dso, ds, dt, offset_s, offset_t, is_arc, xc, yc, zc, mask,
              ia_skip=1, dtype=np.float64, source_zs=None, arc_t_exact=False,
              source_offset=0.0):
    """``cbct_back_fan!`` (src/fbp3/cbct-back.jl:95-144).  ``source_zs`` (na
    source z positions), ``arc_t_exact`` and ``source_offset`` are the options
    the reference leaves commented out (:36-59, :157, :178-180, :188):
    t' = mag (z - source_zs[ia]); for the arc detector mag = dsd /
    sqrt(r_loop^2 + y_beta^2) in t; r_loop = x_beta - source_offset / ds in s'
    and w2 (the flat detector takes it as fbp_back_fan_xy does,
    src/fbp2/back-fan.jl:303, 309-311)."""
    proj = _F(proj, dtype)
    ns, nt, na = proj.shape
    xc = _F(xc, np.float64); yc = _F(yc, np.float64); zc = _F(zc, np.float64)
    mask = _F(mask, np.uint8)
    betas = _F(betas, np.float64)
    img = np.zeros((len(xc), len(yc), len(zc)), dtype=dtype, order="F")
    if source_zs is None and not arc_t_exact and source_offset == 0.0:
        getattr(lib(), "orc_cbct_back_" + _suffix(dtype))(
            _p(proj), ns, nt, na, ia_skip, _p(betas), dsd, dso, ds, dt, offset_s, offset_t,
            int(bool(is_arc)), _p(xc), len(xc), _p(yc), len(yc), _p(zc), len(zc), _p(mask),
            _p(img))
        return img
    zs = None if source_zs is None else _F(source_zs, np.float64)
    if zs is not None and zs.shape != (na,):
        raise ValueError("source_zs needs one z per view")
    getattr(lib(), "orc_cbct_back_opts2_" + _suffix(dtype))(
        _p(proj), ns, nt, na, ia_skip, _p(betas), dsd, dso, ds, dt, offset_s, offset_t,
        int(bool(is_arc)), _p(xc), len(xc), _p(yc), len(yc), _p(zc), len(zc), _p(mask),
        None if zs is None else _p(zs), int(bool(arc_t_exact)), float(source_offset), _p(img))
    return img


def cbct_back_pairs(a, d, betas, dsd, dso, ds, dt, is_arc, xc, yc, zc, mask, dtype=np.float64):
    """``cbct_back_fan!`` (src/fbp3/cbct-back.jl:95-213, offsets 0) on data kept as
    (value, difference) pairs: the bilinear taps of row it+1 are a[., it] + d[., it]
    instead of a[., it+1] (checker for backprojectors that keep the data as
    (value, row difference) pairs; equal to ``cbct_back(a)`` when d is the
    exact row difference)."""
    a = _F(a, dtype)
    d = _F(d, dtype)
    assert a.shape == d.shape
    ns, nt, na = a.shape
    xc = _F(xc, np.float64); yc = _F(yc, np.float64); zc = _F(zc, np.float64)
    mask = _F(mask, np.uint8)
    betas = _F(betas, np.float64)
    img = np.zeros((len(xc), len(yc), len(zc)), dtype=dtype, order="F")
    getattr(lib(), "orc_cbct_back_rp_" + _suffix(dtype))(
        _p(a), _p(d), ns, nt, na, _p(betas), dsd, dso, ds, dt, int(bool(is_arc)),
        _p(xc), len(xc), _p(yc), len(yc), _p(zc), len(zc), _p(mask), _p(img))
    return img


def sino_filter(sino, Hk, extra=0, dtype=np.float64):
    """``fbp_sino_filter`` (src/fbp/filter.jl:122-147): filters dim 1."""
    sino = _F(sino, dtype)
    npad = len(Hk)
    nb = sino.shape[0]
    if nb + extra > npad:
        raise ValueError(f"nb={nb} + extra={extra} > npad={npad}")
    rest = sino.shape[1:]
    ncols = int(np.prod(rest)) if rest else 1
    Hk = np.ascontiguousarray(Hk, dtype=dtype)
    out = np.zeros((nb + extra,) + rest, dtype=dtype, order="F")
    getattr(lib(), "orc_sino_filter_" + _suffix(dtype))(
        _p(sino), nb, ncols, _p(Hk), npad, extra, _p(out))
    return out


def project_bdd(phantom, DSD, DS0, pSize, dSize, nPix, nDet, angle):
    """``project_bdd`` (src/sys2/bdd_2d.jl:17-110): (nangle, nDet)."""
    ph = _F(phantom, np.float64)
    ang = np.ascontiguousarray(angle, dtype=np.float64)
    out = np.zeros((len(ang), nDet), order="F")
    lib().orc_project_bdd(_p(ph), nPix, nDet, _p(ang), len(ang), DSD, DS0, pSize, dSize, _p(out))
    return out


def backproject_bdd(sino, DSD, DS0, pSize, dSize, nPix, nDet, angle, ref_alias=True):
    """``backproject_bdd`` (src/sys2/bdd_2d.jl:119-220): (nPix, nPix)."""
    sn = _F(sino, np.float64)
    ang = np.ascontiguousarray(angle, dtype=np.float64)
    out = np.zeros((nPix, nPix), order="F")
    lib().orc_backproject_bdd(_p(sn), nPix, nDet, _p(ang), len(ang), DSD, DS0, pSize, dSize,
                              int(bool(ref_alias)), _p(out))
    return out


# ---------------------------------------------------------------------------
# whole-path restatements (src/fbp2/fbp.jl:24-86, src/fbp3/fdk.jl:65-81)
# ---------------------------------------------------------------------------

def fbp_par(sino, Hk, view_weight, ar, d, offset, xc, yc, mask, dtype=np.float64):
    """``fbp(plan::FBPNormalPlan{<:SinoPar}, sino)`` (fbp.jl:24-29,72-86)."""
    s = _F(sino, dtype) * np.asarray(view_weight, dtype=dtype)
    f = sino_filter(s, np.asarray(Hk, dtype=dtype), dtype=dtype)
    a = np.asarray(ar, dtype=dtype)
    return back_par(f, np.sin(a), np.cos(a), d, offset, xc, yc, mask, dtype=dtype)


def fbp_fan(sino, Hk, view_weight, sino_weight, ar, dsd, dso, source_offset, is_arc, d,
            offset, xc, yc, mask, dtype=np.float64):
    """``fbp(plan::FBPNormalPlan{<:SinoFan}, sino)`` (fbp.jl:32-38,72-86)."""
    s = _F(sino, dtype) * np.asarray(sino_weight, dtype=dtype)[:, None]
    s = s * np.asarray(view_weight, dtype=dtype)
    f = sino_filter(s, np.asarray(Hk, dtype=dtype), dtype=dtype)
    a = np.asarray(ar, dtype=dtype)
    return back_fan(f, np.sin(a), np.cos(a), dsd, dso, source_offset, is_arc, d, offset,
                    xc, yc, mask, dtype=dtype)


def fdk(proj, Hk, view_weight, ar, dsd, dso, ds, dt, offset_s, offset_t, is_arc,
        xc, yc, zc, mask, dtype=np.float64, **back_options):
    """``fdk(plan, proj)`` (src/fbp3/fdk.jl:65-81); ``back_options`` are the
    backprojector's (``source_zs``, ``arc_t_exact``)."""
    p = _F(proj, dtype) * np.asarray(view_weight, dtype=dtype)
    f = sino_filter(p, np.asarray(Hk, dtype=dtype), dtype=dtype)
    return cbct_back(f, ar, dsd, dso, ds, dt, offset_s, offset_t, is_arc, xc, yc, zc, mask,
                     dtype=dtype, **back_options)


# ---------------------------------------------------------------------------
# the :mojette FBP of a SinoMoj geometry
# ---------------------------------------------------------------------------

def d_moj(d, ar):
    """``_d_moj(rg)(ar)`` = d max(|cos|, |sin|) (src/geom/common.jl:128)."""
    ar = np.asarray(ar, np.float64)
    return d * np.maximum(np.abs(np.cos(ar)), np.abs(np.sin(ar)))


def fbp_moj(sino, ar, d, offset, xc, yc, mask, npad, window=None, decon1=True):
    """Float64 restatement of the :mojette FBP of a SinoMoj sinogram (nb, na):
    each view is ramp-filtered with the response built for its own radial
    spacing dr = _d_moj(ar) (ramp_flat(N, dr), ramp.jl:26-29, through
    fbp_filter, filter.jl:32-57; the per-view filter of the reference's
    commented fbp2_make_sino_filter_moj, fbp.jl:150-170), weighted by the
    angular step (_angle_weights, d-angle.jl:55-65; the "dphi" of fbp.jl:138-145)
    and backprojected pixel by pixel at tau = (x cos + y sin) / dr (tau.jl:22-25)
    with the linear interpolation and bounds of fbp_back_par_xy
    (back-par.jl:292-312)."""
    sino = np.asarray(sino, np.float64)
    nb, na = sino.shape
    ar = np.asarray(ar, np.float64)
    dr = d_moj(d, ar)
    aw = float(angle_weight(np.asarray(ar, np.float32)))
    wb = (nb + 1) / 2 + float(np.float32(offset))
    x = np.asarray(xc, np.float64)[:, None]
    y = np.asarray(yc, np.float64)[None, :]
    img = np.zeros((x.shape[0], y.shape[1]))
    for j in range(na):
        Hk = fbp_filter(ramp_flat(npad, dr[j])[0], npad, dr[j], window, decon1)
        q = sino_filter(sino[:, j:j + 1] * aw, Hk)[:, 0]
        bb = (x * np.cos(ar[j]) + y * np.sin(ar[j])) / dr[j] + wb
        il = np.floor(bb).astype(np.int64)
        ok = (il >= 1) & (il + 1 <= nb)
        ilc = np.clip(il, 1, nb - 1)
        wr = bb - il
        img += np.where(ok, (1 - wr) * q[ilc - 1] + wr * q[ilc], 0.0)
    return np.where(np.asarray(mask, bool), img, 0.0)
