"""Analytic phantoms and their exact projections (synthetic test data).

Random ellipse / ellipsoid phantoms and their line integrals in Float64,
following the reference's ray conventions:
  parallel rays  r = x cos(phi) + y sin(phi)               (src/geom/tau.jl:8-11)
  fan rays       (r, phi) = (dso sin(gamma), gamma + beta)  (src/geom/sino-geom.jl:118-123)
  cone rays      source (-dso sin b, dso cos b, 0) to the rotated detector
                 pixel: flat (s, -dod, t); arc (dsd sin g, dso - dsd cos g, t),
                 g = s / dsd               (src/fbp3/cb_flat_to_par.jl, cb_arc_to_par.jl)
Everything is computed in float64 and returned as column-major (Julia-order)
torch tensors on the requested device: on a GPU the projections run in
libsgx's analytic projector (sgx_radon_ellipses / sgx_radon_ellipsoids), on
the CPU in torch.
"""

from __future__ import annotations

import math

import numpy as np

from .geom import CtFan, SinoFan, SinoMoj, SinoPar


def _torch():
    import torch
    return torch


def _gpu(device):
    """The CUDA device `device` names, or None for the CPU."""
    torch = _torch()
    d = torch.device(device)
    return d if d.type == "cuda" else None


def _colmajor_f64(shape, dev):
    torch = _torch()
    return torch.empty(tuple(reversed(shape)), dtype=torch.float64, device=dev).permute(
        tuple(range(len(shape) - 1, -1, -1)))


def _radon_hip(fn, dev, *args):
    """Call an sgx_radon_* entry point on the device's shared context."""
    from . import _lib
    ctx = _lib.context(dev.index if dev.index is not None else 0)
    with ctx.ordered():
        ctx.check(getattr(_lib.load(), fn)(ctx.h, *args))


def _p(a):
    import ctypes
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def random_ellipses(seed: int, n: int, fov: float, center: float = 0.4,
                    axes=(0.05, 0.4), rho=(-0.5, 1.0)) -> np.ndarray:
    """(n, 6) rows (cx, cy, a, b, theta, rho): centres U(+-center*fov),
    semi-axes U[axes]*fov, rotation U[0, pi), density U[rho]."""
    g = np.random.default_rng(seed)
    out = np.empty((n, 6))
    out[:, 0:2] = g.uniform(-center * fov, center * fov, (n, 2))
    out[:, 2:4] = g.uniform(axes[0] * fov, axes[1] * fov, (n, 2))
    out[:, 4] = g.uniform(0, math.pi, n)
    out[:, 5] = g.uniform(rho[0], rho[1], n)
    return out


def random_ellipsoids(seed: int, n: int, fov: float, center: float = 0.35,
                      axes=(0.05, 0.35), rho=(-0.3, 1.0)) -> np.ndarray:
    """(n, 7) rows (cx, cy, cz, a, b, c, rho), axis-aligned."""
    g = np.random.default_rng(seed)
    out = np.empty((n, 7))
    out[:, 0:3] = g.uniform(-center * fov, center * fov, (n, 3))
    out[:, 3:6] = g.uniform(axes[0] * fov, axes[1] * fov, (n, 3))
    out[:, 6] = g.uniform(rho[0], rho[1], n)
    return out


def ellipse_par(ell, r, phi, device="cpu"):
    """Line integrals of ellipses at parallel rays (r[i], phi[j]) -> (len r, len phi)."""
    torch = _torch()
    r = torch.as_tensor(np.asarray(r, np.float64), device=device)[:, None]
    phi = torch.as_tensor(np.asarray(phi, np.float64), device=device)[None, :]
    return _ellipse_rays(ell, r, phi)


def _ellipse_rays(ell, r, phi):
    torch = _torch()
    out = torch.zeros(torch.broadcast_shapes(r.shape, phi.shape), dtype=torch.float64,
                      device=r.device)
    c, s = torch.cos(phi), torch.sin(phi)
    for cx, cy, a, b, th, rho in np.asarray(ell, np.float64):
        tau = r - (cx * c + cy * s)
        d = phi - th
        s2 = (a * torch.cos(d)) ** 2 + (b * torch.sin(d)) ** 2
        inside = (s2 - tau * tau).clamp_min(0)
        out += 2 * rho * a * b * torch.sqrt(inside) / s2
    return out


def sino_ellipses(ell, rg, device="cpu"):
    """Exact sinogram (nb, na) of ellipses for SinoPar, SinoMoj or SinoFan
    (source_offset 0)."""
    torch = _torch()
    dev = _gpu(device)
    if dev is not None and not isinstance(rg, SinoMoj):
        from . import _lib
        if isinstance(rg, SinoFan) and rg.source_offset != 0:
            raise ValueError("sino_ellipses: source_offset must be 0")
        kind = (_lib.SGX_SINO_PAR if isinstance(rg, SinoPar) else
                _lib.SGX_SINO_FAN_ARC if rg.is_arc else _lib.SGX_SINO_FAN_FLAT)
        e = np.ascontiguousarray(np.asarray(ell, np.float64).reshape(-1, 6))
        sv, bv = np.ascontiguousarray(rg.s()), np.ascontiguousarray(rg.ar())
        out = _colmajor_f64((rg.nb, rg.na), dev)
        fan = isinstance(rg, SinoFan)
        _radon_hip("sgx_radon_ellipses", dev, kind, _p(sv), rg.nb, _p(bv), rg.na,
                   rg.dsd if fan else 0.0, rg.dso if fan else 0.0, _p(e), len(e),
                   out.data_ptr(), 0)
        return out
    return _sino_ellipses_torch(ell, rg, device)


def _sino_ellipses_torch(ell, rg, device):
    """sino_ellipses as torch elementwise operations on any device."""
    torch = _torch()
    s = torch.as_tensor(rg.s(), device=device)[:, None]
    beta = torch.as_tensor(rg.ar(), device=device)[None, :]
    if isinstance(rg, SinoPar):
        return _ellipse_rays(ell, s, beta)
    if isinstance(rg, SinoMoj):
        # the mojette rays (sino-geom.jl:107-115): r = s max(|sin|, |cos|)
        return _ellipse_rays(ell, s * torch.maximum(beta.cos().abs(), beta.sin().abs()), beta)
    if isinstance(rg, SinoFan):
        gam = s / rg.dsd if rg.is_arc else torch.atan(s / rg.dsd)
        return _ellipse_rays(ell, rg.dso * torch.sin(gam), gam + beta)
    raise TypeError(type(rg))


def image_ellipses(ell, x, y, device="cpu"):
    """Pixel-centre samples of the ellipse phantom on the grid x (nx) by y (ny)."""
    torch = _torch()
    X = torch.as_tensor(np.asarray(x, np.float64), device=device)[:, None]
    Y = torch.as_tensor(np.asarray(y, np.float64), device=device)[None, :]
    out = torch.zeros(X.shape[0], Y.shape[1], dtype=torch.float64, device=device)
    for cx, cy, a, b, th, rho in np.asarray(ell, np.float64):
        u = (X - cx) * math.cos(th) + (Y - cy) * math.sin(th)
        v = -(X - cx) * math.sin(th) + (Y - cy) * math.cos(th)
        out += rho * (((u / a) ** 2 + (v / b) ** 2) <= 1)
    return out


def proj_ellipsoids(ell, rg: CtFan, device="cpu", views=None, chunk: int = 64,
                    source_zs=None):
    """Exact cone-beam projections (ns, nt, na) of axis-aligned ellipsoids,
    as a column-major float64 tensor.  ``views`` selects a subset of views;
    ``source_zs`` (na) raises the source and the detector of every view (a
    helical trajectory, ``rg.source_zs()``).  On the CPU, views are processed
    in chunks (one batched pass per ellipsoid)."""
    torch = _torch()
    idx = np.arange(rg.na) if views is None else np.asarray(views)
    zs_all = None if source_zs is None else np.asarray(source_zs, np.float64)
    if zs_all is not None and zs_all.shape != (rg.na,):
        raise ValueError("source_zs needs one z per view")
    dev = _gpu(device)
    if dev is not None:
        from . import _lib
        kind = _lib.SGX_CT_FAN_ARC if rg.is_arc else _lib.SGX_CT_FAN_FLAT
        e = np.ascontiguousarray(np.asarray(ell, np.float64).reshape(-1, 7))
        sv, tv = np.ascontiguousarray(rg.s()), np.ascontiguousarray(rg.t())
        bv = np.ascontiguousarray(np.asarray(rg.ar(), np.float64)[idx])
        zv = None if zs_all is None else np.ascontiguousarray(zs_all[idx])
        out = _colmajor_f64((rg.ns, rg.nt, len(idx)), dev)
        _radon_hip("sgx_radon_ellipsoids", dev, kind, _p(sv), rg.ns, _p(tv), rg.nt, _p(bv),
                   _p(zv), len(idx), rg.dsd, rg.dso, _p(e), len(e), out.data_ptr(), 0)
        return out
    return _proj_ellipsoids_torch(ell, rg, device, idx, zs_all, chunk)


def _proj_ellipsoids_torch(ell, rg, device, idx, zs_all, chunk):
    """proj_ellipsoids as batched torch operations on any device."""
    torch = _torch()
    s = torch.as_tensor(rg.s(), device=device)
    t = torch.as_tensor(rg.t(), device=device)
    beta_all = np.asarray(rg.ar(), np.float64)
    out = torch.zeros((len(idx), rg.nt, rg.ns), dtype=torch.float64, device=device)
    S, T = s[None, None, :], t[None, :, None]          # (1, 1, ns), (1, nt, 1)
    if rg.is_arc:
        g = S / rg.dsd
        lx, ly = rg.dsd * torch.sin(g), rg.dso - rg.dsd * torch.cos(g)
    else:
        lx, ly = S, torch.full_like(S, -rg.dod)
    ell = np.asarray(ell, np.float64)
    for k0 in range(0, len(idx), chunk):
        b = torch.as_tensor(beta_all[idx[k0:k0 + chunk]], device=device)[:, None, None]
        cb, sb = torch.cos(b), torch.sin(b)
        px, py = -rg.dso * sb, rg.dso * cb             # source, (v, 1, 1)
        pz = 0.0 if zs_all is None else \
            torch.as_tensor(zs_all[idx[k0:k0 + chunk]], device=device)[:, None, None]
        dx = (lx * cb - ly * sb) - px                  # (v, 1, ns)
        dy = (lx * sb + ly * cb) - py
        dz = T                                         # (1, nt, 1)
        norm = torch.sqrt(dx * dx + dy * dy + dz * dz)
        ux, uy, uz = dx / norm, dy / norm, dz / norm   # (v, nt, ns)
        acc = out[k0:k0 + chunk]
        for cx, cy, cz, a, bb, c, rho in ell:
            qx, qy, qz = (px - cx) / a, (py - cy) / bb, (pz - cz) / c
            vx, vy, vz = ux / a, uy / bb, uz / c
            A = vx * vx + vy * vy + vz * vz
            B = 2 * (qx * vx + qy * vy + qz * vz)
            C = qx * qx + qy * qy + qz * qz - 1
            disc = (B * B - 4 * A * C).clamp_min(0)
            acc += rho * torch.sqrt(disc) / A
    return out.permute(2, 1, 0)                        # (ns, nt, nv) column-major


def volume_ellipsoids(ell, x, y, z, device="cpu"):
    """Voxel-centre samples of the ellipsoid phantom, (nx, ny, nz) column-major."""
    torch = _torch()
    X = torch.as_tensor(np.asarray(x, np.float64), device=device)[None, None, :]
    Y = torch.as_tensor(np.asarray(y, np.float64), device=device)[None, :, None]
    Z = torch.as_tensor(np.asarray(z, np.float64), device=device)[:, None, None]
    out = torch.zeros(len(z), len(y), len(x), dtype=torch.float64, device=device)
    for cx, cy, cz, a, b, c, rho in np.asarray(ell, np.float64):
        out += rho * ((((X - cx) / a) ** 2 + ((Y - cy) / b) ** 2 + ((Z - cz) / c) ** 2) <= 1)
    return out.permute(2, 1, 0)


def poisson_log(p, photons: float, seed: int):
    """-log(max(k, 1) / I0) with k ~ Poisson(I0 exp(-p)); the zero-count
    clamp keeps the log finite."""
    g = np.random.default_rng(seed)
    p = np.asarray(p, np.float64)
    k = g.poisson(photons * np.exp(-p))
    return -np.log(np.maximum(k, 1) / photons)
