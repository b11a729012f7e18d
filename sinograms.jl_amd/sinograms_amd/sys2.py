"""Distance-driven fan-beam projector (src/sys2/bdd_2d.jl) on libsgx.

``geo`` is the reference's named tuple as a dict or object with DSD, DS0,
pSize, dSize, nPix, nDet and angle (radians).  The phantom is (nPix, nPix);
the sinogram is (length(angle), nDet), Float64, as in the reference.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .recon import _Arr, _ctx_for, _dptr, _empty_colmajor, _f64


def _g(geo, k):
    return geo[k] if isinstance(geo, dict) else getattr(geo, k)


def project_bdd(phantom, geo):
    """``project_bdd(phantom, geo)`` (bdd_2d.jl:17-110), computed in FP64."""
    nPix, nDet = int(_g(geo, "nPix")), int(_g(geo, "nDet"))
    ang = _f64(_g(geo, "angle"))
    is32 = (getattr(phantom, "dtype", None) in (np.float32,)) or \
        str(getattr(phantom, "dtype", "")) == "torch.float32"
    a = _Arr(phantom, dtype=np.float32 if is32 else np.float64)
    if a.shape != (nPix, nPix):
        raise ValueError("phantom size")
    if a.host:
        out = np.zeros((len(ang), nDet), dtype=np.float64, order="F")
        optr = out.ctypes.data
    else:
        import torch
        out = _empty_colmajor((len(ang), nDet), a.device, torch.float64)
        optr = out.data_ptr()
    ctx = _ctx_for(a)
    flags = a.flags | (_lib.SGX_IN_F32 if is32 else 0)
    with ctx.ordered(not a.host):
        ctx.check(_lib.load().sgx_project_bdd(
            ctx.h, a.ptr, nPix, nDet, _dptr(ang), len(ang), float(_g(geo, "DSD")),
            float(_g(geo, "DS0")), float(_g(geo, "pSize")), float(_g(geo, "dSize")),
            ctypes.c_void_p(optr), flags))
    return out


def backproject_bdd(sinogram, geo, ref_alias: bool = True):
    """``backproject_bdd(sinogram, geo)`` (bdd_2d.jl:119-220), FP64.

    ref_alias=True reproduces the reference exactly, including its first
    view being added twice (bdd_2d.jl:139-140); False is the exact adjoint
    of project_bdd."""
    nPix, nDet = int(_g(geo, "nPix")), int(_g(geo, "nDet"))
    ang = _f64(_g(geo, "angle"))
    a = _Arr(sinogram, dtype=np.float64)
    if a.shape != (len(ang), nDet):
        raise ValueError("sinogram size")
    if a.host:
        out = np.zeros((nPix, nPix), dtype=np.float64, order="F")
        optr = out.ctypes.data
    else:
        import torch
        out = _empty_colmajor((nPix, nPix), a.device, torch.float64)
        optr = out.data_ptr()
    ctx = _ctx_for(a)
    with ctx.ordered(not a.host):
        ctx.check(_lib.load().sgx_backproject_bdd(
            ctx.h, a.ptr, nPix, nDet, _dptr(ang), len(ang), float(_g(geo, "DSD")),
            float(_g(geo, "DS0")), float(_g(geo, "pSize")), float(_g(geo, "dSize")),
            int(bool(ref_alias)), ctypes.c_void_p(optr), a.flags))
    return out
