"""Ray conventions of the geometries and analytic projections along them.

Mirrors the reference's ray helpers (host side, float64, vectorised):

  rays(SinoPar)    (r, phi) = (s_i, beta_j)                  src/geom/sino-geom.jl:99-104
  rays(SinoMoj)    (r, phi) = (s_i max|sincos beta_j|, beta_j) src/geom/sino-geom.jl:107-115
  rays(SinoFan)    (dso sin(gamma), gamma + beta)            src/geom/sino-geom.jl:118-132
  rays(CtFan)      (u, v, phi, theta) = cb_*_to_par(s, t, beta, dso, dod)
                                                             src/geom/ct-geom.jl:98-110
  cb_flat_to_par   src/fbp3/cb_flat_to_par.jl:23-30
  cb_arc_to_par    src/fbp3/cb_arc_to_par.jl:23-43
  tau              projected s / ds of pixel centres         src/geom/tau.jl:8-80
  footprint_size   maximum footprint in detector pixels      src/sys2/footprint.jl:8-37

Where the reference returns iterators of tuples, these return the tuple's
arrays, shaped like the sinogram (nb, na) or the projections (ns, nt, na).

A 3-D parallel ray (u, v, phi, theta) is the line u a + v b + l e, l real,
with the orthonormal frame
  a = (cos phi, sin phi, 0),
  b = (sin phi sin theta, -cos phi sin theta, cos theta),
  e = (-sin phi cos theta, cos phi cos theta, sin theta):
cb_arc_to_par's u and v are the source position's components along a and b
and e is its source-to-detector direction.  cb_flat_to_par's polar angle has
the other sign: under the same frame its ray crosses the plane through the
axis perpendicular to the central ray at the point the true ray does, with
the opposite tilt (the reference's "note sign!").  These functions keep both
formulas as the reference has them.
"""

from __future__ import annotations

import math

import numpy as np

from .geom import CtFan, CtSourceCircle, ImageGeom, SinoFan, SinoGeom, SinoMoj, SinoPar


def cb_flat_to_par(s, t, beta, dso: float, dod: float):
    """``cb_flat_to_par(s, t, beta, dso, dod)`` -> (u, v, azim + beta, polar)."""
    s, t, beta = (np.asarray(a, np.float64) for a in (s, t, beta))
    dsd = dso + dod
    hs = np.sqrt(s * s + dsd * dsd)
    u = dso * s / hs
    v = dso * t / np.sqrt(s * s + t * t + dsd * dsd) * dsd / hs
    azim = np.arctan(s / dsd)
    polar = np.arctan(t / hs)
    return u, v, azim + beta, polar


def cb_arc_to_par(s, t, beta, dso: float, dod: float):
    """``cb_arc_to_par(s, t, beta, dso, dod)`` -> (u, v, azim + beta, polar)."""
    s, t, beta = (np.asarray(a, np.float64) for a in (s, t, beta))
    dsd = dso + dod
    rf = dsd                                   # focal distance, dfs = 0
    det = (rf * np.sin(s / rf), dso - rf * np.cos(s / rf), t)
    src = (0.0, dso, 0.0)
    e = [d - c for d, c in zip(det, src)]
    n = np.sqrt(e[0] ** 2 + e[1] ** 2 + e[2] ** 2)
    e = [x / n for x in e]
    azim = -np.arctan2(e[0], e[1])
    polar = np.arcsin(e[2])
    sp, cp = np.sin(azim), np.cos(azim)
    u = cp * src[0] + sp * src[1]
    v = (sp * src[0] - cp * src[1]) * np.sin(polar) + src[2] * np.cos(polar)
    return u, v, azim + beta, polar


def rays(rg):
    """``rays(rg)``: (r, phi) arrays of shape (nb, na) for a 2-D geometry,
    (u, v, phi, theta) arrays of shape (ns, nt, na) for a cone-beam one."""
    if isinstance(rg, SinoPar):
        s, b = np.meshgrid(rg.s(), rg.ar(), indexing="ij")
        return s, b
    if isinstance(rg, SinoMoj):
        # _moj_to_par (sino-geom.jl:107-115): r = s max(|sin|, |cos|)
        s, b = np.meshgrid(rg.s(), rg.ar(), indexing="ij")
        return s * np.maximum(np.abs(np.sin(b)), np.abs(np.cos(b))), b
    if isinstance(rg, SinoFan):
        s, b = np.meshgrid(rg.s(), rg.ar(), indexing="ij")
        g = rg.gamma(s)
        return rg.dso * np.sin(g), g + b
    if isinstance(rg, CtFan):
        if not isinstance(rg.src, CtSourceCircle):
            raise ValueError("non-circular not done")          # ct-geom.jl:99
        s, t, b = np.meshgrid(rg.s(), rg.t(), rg.ar(), indexing="ij")
        f = cb_arc_to_par if rg.is_arc else cb_flat_to_par
        return f(s, t, b, rg.dso, rg.dod)
    raise TypeError(type(rg))


def tau(rg, x=None, y=None):
    """``_tau(rg, x, y)``: projected s / ds of the points (x[i], y[i]) in every
    view, a (len(x), na) matrix (bins from the detector centre); ``tau(rg)``
    is the function (x, y) -> tau(rg, x, y)."""
    if x is None and y is None:
        return lambda x, y: tau(rg, x, y)
    x = np.asarray(x, np.float64).ravel()
    y = np.asarray(y, np.float64).ravel()
    if x.shape != y.shape:
        raise ValueError("x,y axes")
    b = np.asarray(rg.ar(), np.float64)[None, :]
    x, y = x[:, None], y[:, None]
    if isinstance(rg, SinoPar):
        return (x * np.cos(b) + y * np.sin(b)) / rg.d
    if isinstance(rg, SinoMoj):
        # tau.jl:22-25: divided by the view's radial spacing
        return (x * np.cos(b) + y * np.sin(b)) / rg.d_moj(b)
    if isinstance(rg, (SinoFan, CtFan)):
        xb = x * np.cos(b) + y * np.sin(b)
        yb = -x * np.sin(b) + y * np.cos(b)
        if isinstance(rg, SinoFan):
            xb = xb - rg.source_offset
        tangam = xb / (rg.dso - yb)
        ds = rg.d if isinstance(rg, SinoFan) else rg.ds
        return rg.dsd / ds * (np.arctan(tangam) if rg.is_arc else tangam)
    raise TypeError(type(rg))


def footprint_size(rg, ig: ImageGeom) -> np.float32:
    """``footprint_size(rg, ig)``: the maximum footprint of a pixel in
    detector pixels (Float32)."""
    di = math.sqrt(sum(d * d for d in ig.deltas[:2]))
    ds = rg.d if isinstance(rg, SinoGeom) else rg.ds
    if isinstance(rg, SinoPar):
        return np.float32(di / ds)
    if isinstance(rg, (SinoFan, CtFan)):
        rfov = max(abs(n * d) for n, d in zip(ig.dims[:2], ig.deltas[:2])) / 2
        if rfov > 0.99 * rg.dso:
            raise ValueError("bad dso")
        if rg.is_arc:
            return np.float32(di / ds * rg.dsd / (rg.dso - rfov))
        smax = float(np.max(np.abs(rg.s())))
        return np.float32(di / ds * math.sqrt(rg.dsd ** 2 + smax ** 2) / (rg.dso - rfov))
    raise TypeError(type(rg))


def _torch():
    import torch
    return torch


def radon_ellipses(r, phi, ell, device="cpu"):
    """Line integrals of ellipses (rows cx, cy, a, b, theta, rho) along the
    2-D parallel rays (r, phi) (arrays of one shape), float64 torch tensor."""
    from .phantom import _ellipse_rays
    torch = _torch()
    r = torch.as_tensor(np.asarray(r, np.float64), device=device)
    phi = torch.as_tensor(np.asarray(phi, np.float64), device=device)
    return _ellipse_rays(ell, r, phi)


def radon_ellipsoids(u, v, phi, theta, ell, device="cpu"):
    """Line integrals of axis-aligned ellipsoids (rows cx, cy, cz, a, b, c,
    rho) along the 3-D parallel rays (u, v, phi, theta) of the frame above
    (arrays of one shape), float64 torch tensor."""
    torch = _torch()
    u, v, phi, theta = (torch.as_tensor(np.asarray(a, np.float64), device=device)
                        for a in (u, v, phi, theta))
    sp, cp, st, ct = torch.sin(phi), torch.cos(phi), torch.sin(theta), torch.cos(theta)
    px = u * cp + v * sp * st            # a point of the line: u a + v b
    py = u * sp - v * cp * st
    pz = v * ct
    ex, ey, ez = -sp * ct, cp * ct, st   # its unit direction e
    out = torch.zeros_like(u)
    for cx, cy, cz, a, b, c, rho in np.asarray(ell, np.float64):
        qx, qy, qz = (px - cx) / a, (py - cy) / b, (pz - cz) / c
        wx, wy, wz = ex / a, ey / b, ez / c
        A = wx * wx + wy * wy + wz * wz
        B = 2 * (qx * wx + qy * wy + qz * wz)
        C = qx * qx + qy * qy + qz * qz - 1
        out += rho * torch.sqrt((B * B - 4 * A * C).clamp_min(0)) / A
    return out
