"""The reconstruction API of the reference, on libsgx.

Host-side plan construction (ramp, window, filter response, scalar angular
weight) stays on the host, as in the reference; the per-sample work
(weighting, Parker weights, row filtering, backprojection) runs in libsgx.

  fbp_ramp / ramp_flat / ramp_arc     src/fbp/ramp.jl:24-91
  Window, fbp_window                  src/fbp/window.jl:13-91
  fbp_filter                          src/fbp/filter.jl:32-57
  fbp_sino_filter                     src/fbp/filter.jl:122-147       (libsgx)
  parker_weight                       src/fbp/parker.jl:26-154        (fan: libsgx, FP64)
  _angle_weights                      src/fbp/d-angle.jl:209-220
  fbp_sino_weight                     src/fbp2/sino-weight.jl:12-16
  fdk_weight_cyl                      src/fbp3/fdk.jl:21-45
  plan_fbp, FBPNormalPlan, FDKplan    src/fbp2/plan2.jl:34-145, src/fbp3/plan3.jl:19-80
  fbp                                 src/fbp2/fbp.jl:24-86, src/fbp2/fbp-par.jl:23-89
  fdk                                 src/fbp3/fdk.jl:65-81
  fbp_back                            src/fbp2/back-par.jl:30-49, back-fan.jl:11-32
  cbct_back                           src/fbp3/cbct-back.jl:25-45

Arrays use the reference's index order (sino[ib, ia], proj[is, it, ia],
image[ix, iy, iz]) with column-major memory.  numpy inputs are host arrays
(results come back as numpy); torch CUDA tensors stay on the device
(results are CUDA tensors, column-major views).
"""

from __future__ import annotations

import ctypes
import math
import warnings

import numpy as np

from . import _lib
from .geom import (CtFan, ImageGeom, SinoFan, SinoGeom, SinoMoj, SinoPar, is_full_orbit)


class SizeMismatch(ValueError):
    """The reference's ``throw("size mismatch")`` / ``throw("sino size")``."""


# ---------------------------------------------------------------------------
# arrays
# ---------------------------------------------------------------------------

def _is_cuda(x) -> bool:
    return type(x).__module__.startswith("torch") and getattr(x, "is_cuda", False)


def _colmajor(t):
    """A column-major (Julia-order) torch tensor with the same shape."""
    rev = tuple(range(t.dim() - 1, -1, -1))
    if t.permute(rev).is_contiguous():
        return t
    return t.permute(rev).contiguous().permute(rev)


def _empty_colmajor(shape, device, dtype=None):
    import torch
    dtype = dtype or torch.float32
    rev = tuple(range(len(shape) - 1, -1, -1))
    return torch.empty(tuple(reversed(shape)), device=device, dtype=dtype).permute(rev)


class _Arr:
    """An input array as a pointer plus the flags to pass with it."""

    def __init__(self, x, dtype=np.float32):
        if _is_cuda(x):
            import torch
            tdt = {np.float32: torch.float32, np.float64: torch.float64,
                   np.uint8: torch.uint8}[dtype]
            t = _colmajor(x if x.dtype == tdt else x.to(tdt))
            self.keep, self.ptr, self.device, self.host = t, t.data_ptr(), t.device, False
            self.shape = tuple(t.shape)
        else:
            if type(x).__module__.startswith("torch"):
                x = x.detach().cpu().numpy()
            a = np.asfortranarray(np.asarray(x, dtype=dtype))
            self.keep, self.ptr, self.device, self.host = a, a.ctypes.data, None, True
            self.shape = a.shape

    @property
    def flags(self) -> int:
        return _lib.SGX_HOST if self.host else 0

    def out(self, shape):
        if self.host:
            a = np.zeros(shape, dtype=np.float32, order="F")
            return a, a.ctypes.data
        t = _empty_colmajor(shape, self.device)
        return t, t.data_ptr()


def _ctx_for(arr: _Arr):
    dev = arr.device.index if (arr.device is not None and arr.device.index is not None) else 0
    return _lib.context(dev)


def _dptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


# ---------------------------------------------------------------------------
# ramp, window, filter response (host)
# ---------------------------------------------------------------------------

def ramp_flat(N: int, ds: float):
    """Band-limited ramp samples, flat geometry, n = -N/2..N/2-1 (ramp.jl:233-238)."""
    if N % 2:
        raise ValueError("N must be even")
    n = np.arange(-(N // 2), N // 2)
    h = np.zeros(N)
    h[n == 0] = 0.25 / ds ** 2
    odd = (n % 2) == 1
    h[odd] = -1.0 / (math.pi * n[odd] * ds) ** 2
    return h.astype(np.float32), n


def ramp_arc(N: int, ds: float, dsd: float):
    """Band-limited ramp samples, arc geometry (ramp.jl:214-224)."""
    if N % 2:
        raise ValueError("N must be even")
    if N / 2 * ds / dsd > 0.9 * math.pi / 2:
        raise ValueError(f"angle is {math.degrees(N / 2 * ds / dsd)} degrees: too large; "
                         "physically impossible arc geometry")
    n = np.arange(-(N // 2), N // 2)
    h = np.zeros(N)
    h[n == 0] = 0.25 / ds ** 2
    odd = (n % 2) == 1
    h[odd] = -1.0 / (math.pi * dsd * np.sin(n[odd] * ds / dsd)) ** 2
    return h.astype(np.float32), n


def fbp_ramp(rg, N: int):
    """``fbp_ramp`` (ramp.jl:171-182)."""
    if isinstance(rg, SinoPar):
        return ramp_flat(N, rg.d)
    if isinstance(rg, SinoMoj):
        return ramp_flat(N, rg.d)           # dr = rg.d (ramp.jl:26-29)
    if isinstance(rg, (SinoFan, CtFan)):
        ds = rg.d if isinstance(rg, SinoFan) else rg.ds
        return ramp_arc(N, ds, rg.dsd) if rg.is_arc else ramp_flat(N, ds)
    raise TypeError(type(rg))


class NoWindow: pass
class Boxcar: pass
class Hamming: pass
class Hann: pass


class WindowVect:
    def __init__(self, v):
        self.v = np.asarray(v)


class Window:
    """``Window(shape=NoWindow(), cutoff=1)`` (window.jl:13-26)."""

    def __init__(self, shape=None, cutoff: float = 1.0):
        self.shape = shape if shape is not None else NoWindow()
        self.cutoff = cutoff

    def __repr__(self):
        return f"Window({type(self.shape).__name__}(), {self.cutoff})"


def _window_value(shape, width, n):
    inside = np.abs(n) < width / 2
    if isinstance(shape, NoWindow):
        return np.ones(n.shape)
    if isinstance(shape, Boxcar):
        return inside.astype(np.float64)
    if isinstance(shape, Hamming):
        return (0.54 + 0.46 * np.cos(2 * math.pi * n / width)) * inside
    if isinstance(shape, Hann):
        return 0.5 * (1 + np.cos(2 * math.pi * n / width)) * inside
    raise TypeError(type(shape))


def fbp_window(w: Window, N: int) -> np.ndarray:
    """Apodizing window of length N, fftshift-ed, Float32 (window.jl:60-75)."""
    if isinstance(w.shape, WindowVect):
        if len(w.shape.v) != N:
            raise ValueError("N mismatch")
        return np.fft.fftshift(w.shape.v)
    n = np.arange(N) - N // 2
    win = _window_value(w.shape, w.cutoff * N, n).astype(np.float32)
    return np.fft.fftshift(win)


def fbp_filter(rg, npad: int = None, ds: float = None, decon1: bool = True,
               window: Window = None) -> np.ndarray:
    """Apodized ramp filter frequency response Hk, length npad (filter.jl:32-57)."""
    nb = rg.nb if isinstance(rg, SinoGeom) else rg.ns
    npad = npad or _nextpow2(nb + 1)
    if ds is None:
        ds = rg.d if isinstance(rg, SinoGeom) else rg.ds
    window = window or Window()
    h, nn = fbp_ramp(rg, npad)
    Hk = np.fft.fft(np.fft.fftshift(h.astype(np.float64)))
    imag = np.sqrt(np.sum(Hk.imag ** 2) / np.sum(np.abs(Hk) ** 2))
    if imag > math.sqrt(np.finfo(np.float32).eps) * 100:        # _reale (reale.jl:9-13)
        warnings.warn(f"x not real {imag}")
    Hk = Hk.real.astype(np.float32)
    Hk = Hk * fbp_window(window, npad)
    Hk = np.float32(ds) * Hk
    if decon1:
        Hk = Hk / np.fft.fftshift(np.sinc(nn / npad) ** 2)
    return Hk.astype(np.float32)


def _nextpow2(n: int) -> int:
    return 1 << (int(n) - 1).bit_length()


# ---------------------------------------------------------------------------
# weights
# ---------------------------------------------------------------------------

def _angle_weights(ar) -> np.float32:
    """``_angle_weights`` (d-angle.jl:209-220); ar Float32 radians."""
    ar = np.asarray(ar, dtype=np.float32)
    da = np.diff(ar)
    if not np.allclose(da, da[0], rtol=math.sqrt(np.finfo(np.float32).eps), atol=0):
        raise ValueError("Unequal angles not implemented yet")
    last = ar[-1] - ar[-2]
    nxt = ar[-1] + last
    span = nxt - ar[0]
    n180 = max(int(math.floor(1.00001 * abs(float(span)) / math.pi)), 1)
    return np.float32(abs(da[0] / np.float32(n180)))


def fbp_sino_weight(rg: SinoFan) -> np.ndarray:
    """``fbp_sino_weight`` (sino-weight.jl:12-16)."""
    gam = rg.gamma()
    return (np.abs(rg.dso * np.cos(gam) - rg.source_offset * np.sin(gam)) / rg.dsd).astype(
        np.float32)


def fdk_weight_cyl(rg: CtFan) -> np.ndarray:
    """``fdk_weight_cyl`` (fdk.jl:21-45), (ns, nt)."""
    s, t = rg.s()[:, None], rg.t()[None, :]
    if rg.is_arc:
        w = (rg.dso / rg.dsd) * np.cos(s / rg.dsd) / np.sqrt((t / rg.dsd) ** 2 + 1)
    else:
        w = rg.dso / np.sqrt(s ** 2 + t ** 2 + rg.dsd ** 2)
    return np.asfortranarray(w.astype(np.float32))


def parker_weight_par(orbit: float, ad) -> np.ndarray:
    """``parker_weight_par`` (parker.jl:26-55); ad: degrees from orbit_start."""
    ad = np.asarray(ad, dtype=np.float64)
    da = np.diff(ad)
    if len(da) and not np.allclose(da, da[0], rtol=math.sqrt(np.finfo(np.float64).eps)):
        raise ValueError("Unequal angles not yet supported")
    if is_full_orbit(orbit, 180):
        return np.ones((1, 1), dtype=np.float32)
    if abs(orbit) < 180:
        warnings.warn(f"orbit {orbit} < 180")
        return np.ones((1, 1), dtype=np.float32)
    orbit = abs(orbit)
    if orbit > 360:
        raise ValueError("only 180 ≤ |orbit| ≤ 360 supported for Parker weighting")
    extra = orbit - 180
    wt = np.ones((1, len(ad)), dtype=np.float64)
    a = np.abs(ad)
    ii = a < extra
    wt[0, ii] = np.sin(a[ii] / extra * math.pi / 2) ** 2
    ii = a >= 180
    wt[0, ii] = np.sin((orbit - a[ii]) / extra * math.pi / 2) ** 2
    return wt.astype(np.float32)


def parker_weight_fan_short(nb: int, na: int, ar, gam, gamma_max: float,
                            device: int = 0) -> np.ndarray:
    """``parker_weight_fan_short`` (parker.jl:62-108), evaluated in FP64 by libsgx."""
    ar = _f64(ar)
    gam = _f64(gam)
    if len(ar) != na:
        raise ValueError("na bug")
    out = np.zeros((nb, na), dtype=np.float32, order="F")
    ctx = _lib.context(device)
    ctx.check(_lib.load().sgx_parker_fan_short(
        ctx.h, nb, na, _dptr(ar), _dptr(gam), float(gamma_max), _dptr(out), _lib.SGX_HOST))
    return out


def parker_weight(rg, device: int = 0) -> np.ndarray:
    """``parker_weight`` (parker.jl:58-59, 111-119, 147-154)."""
    if isinstance(rg, SinoPar):
        return parker_weight_par(rg.orbit, rg.angles() - rg.orbit_start)
    if isinstance(rg, SinoMoj):
        # parker.jl:122-129: only whole multiples of 180 degrees
        if math.trunc(rg.orbit / 180) * 180 != abs(rg.orbit):
            raise ValueError(f"No Parker weighting for Mojette geometry with orbit={rg.orbit}")
        return np.ones((1, 1), dtype=np.float32)
    if isinstance(rg, SinoFan):
        if is_full_orbit(rg.orbit, 360):
            return np.ones((1, 1), dtype=np.float32)
        _short_warnings(rg)
        return parker_weight_fan_short(rg.nb, rg.na, rg.ar(f32=True), rg.gamma(), rg.gamma_max(),
                                       device)
    if isinstance(rg, CtFan):
        if is_full_orbit(rg.orbit, 360):
            return np.ones((1, 1, 1), dtype=np.float32)
        _short_warnings(rg)
        w = parker_weight_fan_short(rg.ns, rg.na, rg.ar(f32=True), rg.gamma(), rg.gamma_max(),
                                    device)
        return np.asfortranarray(w.reshape(rg.ns, 1, rg.na, order="F"))
    raise TypeError(type(rg))


def _short_warnings(rg):
    os_ = rg.orbit_short()
    if rg.orbit < os_ - np.finfo(np.float64).eps:
        warnings.warn(f"orbit {rg.orbit} is less than a short scan {os_}")
    ar = rg.ar()
    if rg.na > 1 and rg.orbit > os_ + math.degrees(ar[1] - ar[0]):
        warnings.warn(f"orbit {rg.orbit} exceeds short scan {os_}")


# ---------------------------------------------------------------------------
# row filter
# ---------------------------------------------------------------------------

def fbp_sino_filter(sino, filter, extra: int = 0, precision: str = "fp32"):
    """Ramp-filter every column of ``sino`` (nb, ...) with the response
    ``filter`` (npad,) (filter.jl:122-147).  Returns (nb + extra, ...)."""
    a = _Arr(sino)
    Hk = np.ascontiguousarray(np.asarray(filter, dtype=np.float32))
    npad = len(Hk)
    nb = a.shape[0]
    if nb + extra > npad:
        raise ValueError(f"nb={nb} + extra={extra} > npad={npad}")
    ncols = int(np.prod(a.shape[1:])) if len(a.shape) > 1 else 1
    out, optr = a.out((nb + extra,) + tuple(a.shape[1:]))
    ctx = _ctx_for(a)
    prec = _lib.SGX_PREC_BF16 if precision == "bf16x3" else _lib.SGX_PREC_FP32
    with ctx.ordered(not a.host):
        ctx.check(_lib.load().sgx_sino_filter(ctx.h, a.ptr, nb, ncols, _dptr(Hk), npad, extra,
                                              prec, optr, a.flags))
    return out


# ---------------------------------------------------------------------------
# backprojectors
# ---------------------------------------------------------------------------

def _mask_arr(ig: ImageGeom, like: _Arr):
    m = ig.mask
    if like.host:
        return np.asfortranarray(m.astype(np.uint8))
    import torch
    mt = torch.from_numpy(np.ascontiguousarray(m.astype(np.uint8).T)).to(like.device)
    return mt.permute(tuple(range(mt.dim() - 1, -1, -1)))


def _mask_ptr(m):
    return ctypes.c_void_p(m.ctypes.data) if isinstance(m, np.ndarray) else ctypes.c_void_p(
        m.data_ptr())


def fbp_back(rg, ig: ImageGeom, sino, ia_skip: int = 1):
    """Pixel-driven 2-D backprojection (back-par.jl:30-49, back-fan.jl:11-32)."""
    a = _Arr(sino)
    if a.shape != rg.dims():
        raise SizeMismatch("sino size")
    if len(ig.dims) != 2:
        raise SizeMismatch("size mismatch")
    nb, na = rg.dims()
    ar32 = rg.ar(f32=True)[::ia_skip].astype(np.float32)
    sn = _f64(np.sin(ar32))
    cs = _f64(np.cos(ar32))
    xc, yc = (_f64(v) for v in ig.axes())
    mask = _mask_arr(ig, a)
    img, iptr = a.out(ig.dims)
    ctx = _ctx_for(a)
    L = _lib.load()
    if not isinstance(rg, (SinoPar, SinoFan)):
        raise TypeError(type(rg))
    with ctx.ordered(not a.host):
        if isinstance(rg, SinoPar):
            rc = L.sgx_back_par(ctx.h, a.ptr, nb, na, ia_skip, _dptr(sn), _dptr(cs), rg.d,
                                rg.offset, _dptr(xc), len(xc), _dptr(yc), len(yc),
                                _mask_ptr(mask), iptr, a.flags)
        else:
            rc = L.sgx_back_fan(ctx.h, a.ptr, nb, na, ia_skip, _dptr(sn), _dptr(cs), rg.dsd,
                                rg.dso, rg.source_offset, int(rg.is_arc), rg.d, rg.offset,
                                _dptr(xc), len(xc), _dptr(yc), len(yc), _mask_ptr(mask), iptr,
                                a.flags)
        ctx.check(rc)
    return img


def _cbct_opts(rg: CtFan, use_source_zs: bool, arc_t_exact: bool,
               use_source_offset: bool = False):
    """The cone-beam options (None: the reference's geometry) and the array
    they point to, which must stay alive for the call."""
    so = float(rg.source_offset) if use_source_offset else 0.0
    if not use_source_zs and not arc_t_exact and so == 0.0:
        return None, None
    zs = _f64(rg.source_zs()) if use_source_zs else None
    return _lib.SgxCbctOpts(_dptr(zs), int(bool(arc_t_exact)), so), zs


def cbct_back(proj, rg: CtFan, ig: ImageGeom, ia_skip: int = 1, use_source_zs: bool = False,
              arc_t_exact: bool = False, use_source_offset: bool = False):
    """Voxel-driven cone-beam backprojection (cbct-back.jl:25-45).

    Options the reference keeps commented out (:36-59, :157, :178-180, :188),
    off by default: ``use_source_zs`` backprojects with the source z of every
    view, ``rg.source_zs()`` (a helical ``CtSourceHelix`` trajectory), t' =
    mag (z - zs); ``arc_t_exact`` uses, for the arc detector, the t
    magnification of the ray that meets the detector cylinder,
    dsd / sqrt(r^2 + y_b^2); ``use_source_offset`` applies ``rg.source_offset``,
    r = x_b - source_offset / ds in s' and w2 (for both detectors, as the 2-D
    fan backprojector does, back-fan.jl:303-311).
    """
    a = _Arr(proj)
    if a.shape != rg.dims() or len(ig.dims) != 3:
        raise SizeMismatch("size mismatch")
    ns, nt, na = rg.dims()
    betas = _f64(rg.ar(f32=True))
    xc, yc, zc = (_f64(v) for v in ig.axes())
    mask = _mask_arr(ig, a)
    img, iptr = a.out(ig.dims)
    ctx = _ctx_for(a)
    opts, _keep = _cbct_opts(rg, use_source_zs, arc_t_exact, use_source_offset)
    args = (ctx.h, a.ptr, ns, nt, na, ia_skip, _dptr(betas), rg.dsd, rg.dso, rg.ds, rg.dt,
            rg.offset_s, rg.offset_t, int(rg.is_arc), _dptr(xc), len(xc), _dptr(yc), len(yc),
            _dptr(zc), len(zc), _mask_ptr(mask), iptr, a.flags)
    with ctx.ordered(not a.host):
        if opts is None:
            ctx.check(_lib.load().sgx_cbct_back(*args))
        else:
            ctx.check(_lib.load().sgx_cbct_back_opts(*args, ctypes.byref(opts)))
    return img


# ---------------------------------------------------------------------------
# plans
# ---------------------------------------------------------------------------

class _Plan:
    """A libsgx plan: geometry tables, filter kernel, weights and the
    filtered-data buffer live on the device."""

    def __init__(self, rg, ig: ImageGeom, filter, view_weight, precision: str, device: int):
        self.rg, self.ig, self.filter, self.view_weight = rg, ig, filter, view_weight
        self.precision = precision
        self.device = device
        self._ctx = _lib.context(device)
        G = _lib.SgxGeom()
        if isinstance(rg, SinoPar):
            G.kind = _lib.SGX_SINO_PAR
        elif isinstance(rg, SinoMoj):
            G.kind = _lib.SGX_SINO_MOJ
        elif isinstance(rg, SinoFan):
            G.kind = _lib.SGX_SINO_FAN_ARC if rg.is_arc else _lib.SGX_SINO_FAN_FLAT
        elif isinstance(rg, CtFan):
            G.kind = _lib.SGX_CT_FAN_ARC if rg.is_arc else _lib.SGX_CT_FAN_FLAT
        else:
            raise TypeError(type(rg))
        is3 = isinstance(rg, CtFan)
        fan = isinstance(rg, (SinoFan, CtFan))
        G.ns = rg.ns if is3 else rg.nb
        G.nt = rg.nt if is3 else 1
        G.na = rg.na
        G.ds = rg.ds if is3 else rg.d
        G.dt = rg.dt if is3 else 0.0
        G.offset_s = rg.offset_s if is3 else rg.offset
        G.offset_t = rg.offset_t if is3 else 0.0
        G.dsd = rg.dsd if fan else 0.0
        G.dso = rg.dso if fan else 0.0
        G.source_offset = rg.source_offset if fan else 0.0
        G.gamma_max = rg.gamma_max() if fan else 0.0
        G.parker = int(fan and not is_full_orbit(rg.orbit, 360))
        self.ar = rg.ar(f32=True)
        G.angle_weight = float(_angle_weights(self.ar))
        self.geom = G
        xyz = [_f64(v) for v in ig.axes()]
        if is3 and len(xyz) != 3 or (not is3 and len(xyz) != 2):
            raise SizeMismatch("size mismatch")
        zc = xyz[2] if is3 else None
        mask = None if ig.mask.all() else np.asfortranarray(ig.mask.astype(np.uint8))
        Hk = np.ascontiguousarray(np.asarray(filter, dtype=np.float32))
        prec = {"fp32": _lib.SGX_PREC_FP32, "bf16": _lib.SGX_PREC_BF16,
                "f16s": _lib.SGX_PREC_F16S}[precision]
        h = ctypes.c_void_p()
        self._ctx.check(_lib.load().sgx_plan_create(
            self._ctx.h, ctypes.byref(G), _dptr(_f64(self.ar)), _dptr(xyz[0]), len(xyz[0]),
            _dptr(xyz[1]), len(xyz[1]), _dptr(zc), len(zc) if zc is not None else 1,
            _dptr(mask), _dptr(Hk), len(Hk), prec, ctypes.byref(h)))
        self.h = h
        if isinstance(rg, SinoPar):
            pw = parker_weight_par(rg.orbit, rg.angles() - rg.orbit_start)
            if pw.size > 1:
                w = np.ascontiguousarray(pw.ravel().astype(np.float32))
                self._ctx.check(_lib.load().sgx_plan_set_view_weight(self.h, _dptr(w)))

    @property
    def in_dims(self):
        return self.rg.dims()

    def execute(self, sino, out=None):
        """Weight, filter and backproject ``sino``; ``out`` may be a
        preallocated column-major CUDA tensor of the image size."""
        a = _Arr(sino)
        if a.host and self.device != 0:
            raise ValueError("host arrays run on device 0")
        if out is not None:
            if a.host or tuple(out.shape) != tuple(self.ig.dims) or \
                    not out.permute(*range(out.dim() - 1, -1, -1)).is_contiguous():
                raise ValueError("out must be a column-major device tensor of the image size")
            img, iptr = out, out.data_ptr()
        else:
            img, iptr = a.out(self.ig.dims)
        # device arrays on torch's current stream: stream-ordered, no host
        # wait; on another stream the call waits for that stream's queued work
        with self._ctx.ordered(not a.host) as same:
            flags = a.flags | (_lib.SGX_ASYNC if same else 0)
            self._ctx.check(_lib.load().sgx_plan_execute(self.h, a.ptr, iptr, flags))
        return img

    def times(self):
        """(filter_ms, back_ms) of the last execute (HIP events)."""
        f, b = ctypes.c_float(), ctypes.c_float()
        self._ctx.check(_lib.load().sgx_plan_last_times(self.h, ctypes.byref(f),
                                                        ctypes.byref(b)))
        return f.value, b.value

    def applied_weights(self) -> np.ndarray:
        """All pre-filter weights the plan applies, full (ns, nt, na) size."""
        G = self.geom
        out = np.zeros((G.ns, G.nt, G.na), dtype=np.float32, order="F")
        self._ctx.check(_lib.load().sgx_plan_view_weight(self.h, _dptr(out)))
        return out

    def filtered_ptr(self) -> int:
        return _lib.load().sgx_plan_filtered(self.h)

    def read_filtered(self):
        """(stored filtered data as float32 (ns, nt, na) column-major, per-view
        scale) of the last execute; the data are the stored values (bf16 /
        fp16 widened exactly), the scale multiplies them."""
        G = self.geom
        n = G.ns * G.nt * G.na
        if self.precision == "fp32":
            raw = np.empty(n, np.float32)
        else:
            raw = np.empty(n, np.uint16)
        sc = np.empty(G.na, np.float32)
        self._ctx.check(_lib.load().sgx_plan_read_filtered(self.h, _dptr(raw), _dptr(sc)))
        if self.precision == "bf16":
            val = (raw.astype(np.uint32) << 16).view(np.float32)
        elif self.precision == "f16s":
            val = raw.view(np.float16).astype(np.float32)
        else:
            val = raw
        return val.reshape((G.ns, G.nt, G.na), order="F"), sc

    def read_stored(self):
        """The stored filtered data of the last execute as-is: (pitch, nt, na)
        column-major array of the storage type (float32, or uint16 bit patterns
        of bf16 / fp16), pitch >= ns with zero pad bins."""
        G = self.geom
        L = _lib.load()
        nbytes = L.sgx_plan_stored_bytes(self.h)
        pitch = L.sgx_plan_pitch(self.h)
        raw = np.empty(nbytes, np.uint8)
        self._ctx.check(L.sgx_plan_read_stored(self.h, _dptr(raw), nbytes))
        v = raw.view(np.float32 if self.precision == "fp32" else np.uint16)
        return v.reshape((pitch, G.nt, G.na), order="F")

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.load().sgx_plan_destroy(self.h)
                self.h = None
        except Exception:
            pass


class FBPNormalPlan(_Plan):
    """``FBPNormalPlan`` (plan2.jl:34-45): rg, ig, filter Hk, view_weight."""

    def __repr__(self):
        return (f"FBPNormalPlan({type(self.rg).__name__} {self.rg.dims()}, ig {self.ig.dims}, "
                f"H extrema {self.filter.min():.4g} {self.filter.max():.4g}, "
                f"V {np.shape(self.view_weight)})")


class MojPlan(_Plan):
    """The ``:mojette`` FBP plan of a ``SinoMoj`` geometry (the reference's
    commented ``MojPlan``, plan2.jl:56-62, 125-135, and ``fpb2(plan::MojPlan)``,
    fbp.jl:138-145): every view is ramp-filtered for its own radial spacing
    d_moj = d max(|cos|, |sin|) (the response of spacing d divided by
    max(|cos|, |sin|), fbp.jl:150-160) and backprojected with that spacing,
    times the angular weight dphi."""

    def __repr__(self):
        return (f"MojPlan(SinoMoj {self.rg.dims()}, ig {self.ig.dims}, "
                f"H extrema {self.filter.min():.4g} {self.filter.max():.4g})")


class FDKplan(_Plan):
    """``FDKplan`` (plan3.jl:19-29)."""

    def __repr__(self):
        return (f"FDKplan({type(self.rg).__name__} {self.rg.dims()}, ig {self.ig.dims}, "
                f"precision {self.precision})")


def _fbp_weights(rg, device=0):
    """``_fbp_weights`` = parker_weight .* _angle_weights (plan2.jl:141-145)."""
    return parker_weight(rg, device) * _angle_weights(rg.ar(f32=True))


def _fdk_weights(rg, device=0):
    """``_fdk_weights`` = parker .* fdk_weight_cyl .* _angle_weights (plan3.jl:75-80)."""
    return (parker_weight(rg, device) * fdk_weight_cyl(rg)[:, :, None]
            * _angle_weights(rg.ar(f32=True)))


# plan_fbp's `how` (plan2.jl:93-96, 119-133): the reference builds only
# :normal; :mojette, :dsc, :df and :table are names in a commented dispatch
# onto MIRT objects (Gtomo2_table, Gtomo2_dscex, a NUFFT) that it does not
# contain.  :mojette is built here from the mojette geometry it does define.
_HOW_UNDEFINED = ("dsc", "df", "table")


def plan_fbp(rg, ig: ImageGeom, window: Window = None, npad: int = None, decon1: bool = True,
             filter=None, precision: str = "fp32", device: int = 0,
             use_source_zs: bool = False, arc_t_exact: bool = False,
             use_source_offset: bool = False, how: str = "normal", **_ignored):
    """``plan_fbp`` (plan2.jl:106-138 for SinoGeom, plan3.jl:60-72 for CtGeom).

    how: "normal" (the reference's FBP), or "mojette" for a ``SinoMoj``
    geometry (also its default: the reference has no other FBP for it), the
    per-view-spacing FBP of ``MojPlan``.  "dsc", "df" and "table" raise: the
    reference names them without an algorithm.

    precision: "fp32" (FP32 filter and storage), or for cone-beam plans the
    16-bit storage forms "bf16" (split-BF16x3 MFMA filter, BF16 filtered
    projections) and "f16s" (split-FP16x3 MFMA filter with power-of-two
    scaled operands, FP16 filtered projections scaled per view by a power of
    two); backprojection always accumulates in FP32.

    use_source_zs, arc_t_exact, use_source_offset (cone-beam plans): the
    backprojection options of ``cbct_back``, which the reference keeps
    commented out; off by default.
    """
    how = str(how).lstrip(":")
    if how in _HOW_UNDEFINED:
        raise NotImplementedError(f"how=:{how}: the reference defines no algorithm for it "
                                  "(plan2.jl:125-135 keep it commented out)")
    if how not in ("normal", "mojette"):
        raise ValueError(f"unknown type: {how}")                # plan2.jl:134
    if how == "mojette" and not isinstance(rg, SinoMoj):
        raise ValueError("how=:mojette needs a SinoMoj geometry (the reference's rebinning of "
                         "other geometries onto mojette sampling is not defined)")
    nb = rg.nb if isinstance(rg, SinoGeom) else rg.ns
    npad = npad or _nextpow2(nb + 1)
    if filter is None:
        filter = fbp_filter(rg, npad=npad, window=window or Window(), decon1=decon1)
    if isinstance(rg, SinoMoj):
        parker_weight(rg)                                       # orbit check (parker.jl:122-129)
        if use_source_zs or arc_t_exact or use_source_offset:
            raise ValueError("use_source_zs, arc_t_exact and use_source_offset apply to "
                             "cone-beam geometries")
        return MojPlan(rg, ig, filter, np.ones((1, 1), np.float32), precision, device)
    if isinstance(rg, CtFan):
        vw = np.float32(_angle_weights(rg.ar(f32=True)))   # the full array is on the device
        plan = FDKplan(rg, ig, filter, vw, precision, device)
        opts, _keep = _cbct_opts(rg, use_source_zs, arc_t_exact, use_source_offset)
        if opts is not None:
            plan._ctx.check(_lib.load().sgx_plan_set_cbct_opts(plan.h, ctypes.byref(opts)))
        return plan
    if use_source_zs or arc_t_exact or use_source_offset:
        raise ValueError("use_source_zs, arc_t_exact and use_source_offset apply to cone-beam "
                         "geometries")
    vw = parker_weight_par(rg.orbit, rg.angles() - rg.orbit_start) if isinstance(rg, SinoPar) \
        else np.ones((1, 1), dtype=np.float32)
    return FBPNormalPlan(rg, ig, filter, vw, precision, device)


def fbp(plan_or_sino, sino=None, **kw):
    """``fbp(plan, sino)`` (fbp.jl:24-68) or ``fbp(sino; dr, ...)`` (fbp-par.jl:23-89)."""
    if sino is None:
        return _fbp_par_simple(plan_or_sino, **kw)
    plan = plan_or_sino
    if isinstance(plan, FDKplan):
        return fdk(plan, sino)
    shape = tuple(sino.shape)
    if len(shape) == 2:
        if shape != plan.rg.dims():
            raise ValueError("bad sino size")
        return plan.execute(sino)
    # stack of sinograms: serial slice loop (fbp.jl:46-68)
    nb, na = shape[:2]
    rest = shape[2:]
    L = int(np.prod(rest))
    if _is_cuda(sino):
        # (L, na, nb) row-major view of the column-major stack
        s3 = _colmajor(sino).permute(*range(len(shape) - 1, -1, -1)).reshape(L, na, nb)
        out = _empty_colmajor(plan.ig.dims + (L,), sino.device)
        for l in range(L):
            out[:, :, l] = plan.execute(s3[l].permute(1, 0))
        rev = tuple(range(len(plan.ig.dims) + len(rest) - 1, -1, -1))
        return out.permute(2, 1, 0).reshape(tuple(reversed(rest)) + plan.ig.dims[::-1]).permute(rev)
    s3 = np.asfortranarray(np.asarray(sino, dtype=np.float32)).reshape((nb, na, L), order="F")
    out = np.zeros(plan.ig.dims + (L,), dtype=np.float32, order="F")
    for l in range(L):
        out[:, :, l] = plan.execute(s3[:, :, l])
    return out.reshape(plan.ig.dims + rest, order="F")


def fdk(plan: FDKplan, proj):
    """``fdk(plan, proj)`` (fdk.jl:65-81)."""
    if tuple(proj.shape) != plan.rg.dims():
        raise ValueError(f"size mismatch {tuple(proj.shape)} {plan.rg.dims()}")
    return plan.execute(proj)


def _fbp_par_simple(sino, nx: int = None, ny: int = None, dr: float = 1.0, dx: float = None,
                    dy: float = None, orbit: float = 180.0, orbit_start: float = 0.0,
                    offsets=(0.0, 0.0), rmax: float = 0.0, **kw):
    """``fbp(sino; dr, ...)`` for a parallel-beam sinogram (fbp-par.jl:23-89)."""
    nb, na = tuple(sino.shape)
    nx = nx or nb
    ny = ny or nx
    dx = dx or dr
    dy = dy or dx
    rg = SinoPar(nb=nb, na=na, d=dr, orbit=orbit, orbit_start=orbit_start)
    ig = ImageGeom((nx, ny), (dx, dy), offsets)
    ig = ig.with_mask(ig.circle(rmax if rmax else rg.rfov()))
    plan = plan_fbp(rg, ig, **kw)
    return fbp(plan, sino)
