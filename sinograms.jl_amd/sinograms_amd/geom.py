"""Ray geometries and image grids.

Mirrors the reference's geometry records and the derived quantities the hot
path consumes:
  SinoPar / SinoFanArc / SinoFanFlat   src/geom/type2.jl:95-159, ctors :182-345
  SinoMoj, _d_moj, _d_ang              src/geom/type2.jl:108-121, common.jl:60, 128-129
  CtFanArc / CtFanFlat                 src/geom/type3.jl:145-184, ctors :306-409
  angles, _ar, _dso, _gamma, _rfov ... src/geom/common.jl:27-99
  _s, _w                               src/geom/sino-geom.jl:18-20
  _s, _t, _ws, _wt                     src/geom/ct-geom.jl:19-23
  CtSourceCircle, CtSourceHelix        src/geom/ct-source.jl:22-57
  _cone_angle, _zfov, _source_dz_per_view, _source_zs
                                       src/geom/ct-geom.jl:115-140
The `:short` constructors are spelled ``SinoFanArc.short(...)``.

Scalars are Python floats.  The reference's Float32 geometry records are
reproduced where it matters to the hot path: angles can be produced in
Float32 (``ar(f32=True)``), and the bin offsets pass through Float32
(``Toffset``, src/geom/util.jl:5).
"""

from __future__ import annotations

import dataclasses
import math
from typing import Optional, Tuple

import numpy as np


def _f32(x) -> float:
    return float(np.float32(x))


@dataclasses.dataclass(frozen=True)
class SinoGeom:
    nb: int = 128
    d: float = 1.0
    offset: float = 0.0
    na: int = 0
    orbit: float = 180.0
    orbit_start: float = 0.0

    # common.jl / sino-geom.jl methods
    def dims(self) -> Tuple[int, int]:
        return (self.nb, self.na)

    def angles(self) -> np.ndarray:
        """``angles(rg)``: degrees (common.jl:27-28)."""
        return self.orbit_start + np.arange(self.na) * (self.orbit / self.na)

    def ar(self, f32: bool = False) -> np.ndarray:
        """``_ar(rg)``: radians (common.jl:34, unit.jl:10).  With f32 the values
        are the Float32 ones of a Float32 geometry (range formed in extended
        precision, rounded once)."""
        if f32:
            k = float(np.float32(math.pi / 180))
            deg = _f32(self.orbit_start) + np.arange(self.na) * (_f32(self.orbit) / self.na)
            return (deg * k).astype(np.float32).astype(np.float64)
        return self.angles() * (math.pi / 180)

    def w(self) -> float:
        return (self.nb - 1) / 2 + _f32(self.offset)

    def s(self) -> np.ndarray:
        """``_s(rg)`` (sino-geom.jl:18-20)."""
        return self.d * (np.arange(self.nb) - self.w())

    @property
    def ds(self) -> float:
        return self.d

    def ones(self, dtype=np.float32):
        return np.ones(self.dims(), dtype=dtype, order="F")

    def cone_angle(self) -> float:
        """``_cone_angle`` = atan(nt dt / 2 / dsd) (ct-geom.jl:115)."""
        return math.atan(self.nt * self.dt / 2 / self.dsd)

    def zfov(self) -> float:
        """``_zfov`` = dso / dsd * nt * dt (ct-geom.jl:119)."""
        return self.dso / self.dsd * self.nt * self.dt

    def source_dz_per_view(self) -> float:
        """``_source_dz_per_view`` (ct-geom.jl:122-131): 0 for a circle; for a
        helix pitch * zfov / (na * 360 / orbit), and 0 for one view."""
        if not isinstance(self.src, CtSourceHelix) or self.na == 1:
            return 0.0
        na_per_360 = self.na * (360 / self.orbit)
        return self.src.pitch * self.zfov() / na_per_360

    def source_zs(self) -> np.ndarray:
        """``_source_zs`` (ct-geom.jl:134-140): the source z of every view."""
        if not isinstance(self.src, CtSourceHelix):
            return np.zeros(self.na)
        return np.arange(self.na) * self.source_dz_per_view() + self.src.source_z0

    def zeros(self, dtype=np.float32):
        return np.zeros(self.dims(), dtype=dtype, order="F")


@dataclasses.dataclass(frozen=True)
class SinoPar(SinoGeom):
    """2-D parallel-beam sinogram geometry (type2.jl:95-105, ctor :196-210)."""

    def __post_init__(self):
        if self.na <= 0:
            object.__setattr__(self, "na", 2 * int(math.floor(self.nb * math.pi / 2 / 2)))

    def rfov(self) -> float:
        return float(np.max(np.abs(self.s())))


@dataclasses.dataclass(frozen=True)
class SinoMoj(SinoGeom):
    """2-D mojette sinogram geometry (type2.jl:108-121, ctor :199-229): ``d``
    is the (square) pixel size, and view phi samples the radial axis at
    ``d_moj(phi) = d max(|cos phi|, |sin phi|)`` (``_d_moj``, common.jl:128)."""

    def __post_init__(self):
        if self.na <= 0:
            object.__setattr__(self, "na", 2 * int(math.floor(self.nb * math.pi / 2 / 2)))

    def d_moj(self, ar) -> np.ndarray:
        """``_d_moj(rg)(ar)`` (common.jl:128)."""
        ar = np.asarray(ar, np.float64)
        return self.d * np.maximum(np.abs(np.cos(ar)), np.abs(np.sin(ar)))

    def d_ang(self) -> np.ndarray:
        """``_d_ang(rg)``: the radial spacing of every view (common.jl:129)."""
        return self.d_moj(self.ar())

    def rfov(self) -> float:
        """``_rfov(rg::SinoMoj)`` = nb / 2 min(d_ang) (common.jl:60, ignores offset)."""
        return self.nb / 2 * float(np.min(self.d_ang()))


@dataclasses.dataclass(frozen=True)
class SinoFan(SinoGeom):
    orbit: float = 360.0
    source_offset: float = 0.0
    dsd: float = 0.0
    dod: float = 0.0

    is_arc = True

    def __post_init__(self):
        if self.na <= 0:
            object.__setattr__(self, "na", 2 * int(math.floor(self.nb * math.pi / 2 / 2)))
        if self.dsd == 0.0:
            object.__setattr__(self, "dsd", 4.0 * self.nb * self.d)
        if self.dod == 0.0:
            object.__setattr__(self, "dod", 1.0 * self.nb * self.d)

    @property
    def dso(self) -> float:
        """``_dso`` = dsd - dod (common.jl:53)."""
        return self.dsd - self.dod

    def gamma(self, s=None) -> np.ndarray:
        """``_gamma`` (common.jl:93-97)."""
        s = self.s() if s is None else np.asarray(s, dtype=np.float64)
        return s / self.dsd if self.is_arc else np.arctan(s / self.dsd)

    def gamma_max(self) -> float:
        return float(np.max(np.abs(self.gamma())))

    def orbit_short(self) -> float:
        """``_orbit_short`` in degrees (common.jl:40)."""
        return 180 + 2 * math.degrees(self.gamma_max())

    def rfov(self) -> float:
        """``_rfov`` = dso sin(gamma_max) (common.jl:61)."""
        return self.dso * math.sin(self.gamma_max())

    @classmethod
    def _short(cls, na: int = 128, orbit_start: float = 0.0, **kw):
        # type2.jl:281-288: na scaled by orbit_short / 360, orbit = orbit_short
        tmp = cls(orbit=360.0, **kw)
        os_ = tmp.orbit_short()
        na2 = int(math.ceil(na * os_ / 360))
        return cls(na=na2, orbit=os_, orbit_start=orbit_start, **kw)


@dataclasses.dataclass(frozen=True)
class SinoFanArc(SinoFan):
    """2-D fan-beam, arc detector (type2.jl:127-141, ctor :264-288)."""
    is_arc = True

    @classmethod
    def short(cls, **kw):
        return cls._short(**kw)


@dataclasses.dataclass(frozen=True)
class SinoFanFlat(SinoFan):
    """2-D fan-beam, flat detector (type2.jl:145-159, ctor :321-345)."""
    is_arc = False

    @classmethod
    def short(cls, **kw):
        return cls._short(**kw)


@dataclasses.dataclass(frozen=True)
class CtSourceCircle:
    """Circular source trajectory (ct-source.jl:22-26)."""


@dataclasses.dataclass(frozen=True)
class CtSourceHelix:
    """Helical source trajectory of constant pitch (ct-source.jl:29-57):
    ``pitch`` is a unitless fraction of the z field of view per 360 degrees
    (stored as Float32, as the reference does), ``source_z0`` the source z of
    the first view, in the units of the detector."""
    pitch: float = 0.0
    source_z0: float = 0.0

    def __post_init__(self):
        object.__setattr__(self, "pitch", _f32(self.pitch))


@dataclasses.dataclass(frozen=True)
class CtFan:
    """3-D cone-beam geometry (type3.jl:145-184); ``src`` is its source
    trajectory, circular by default."""
    ns: int = 128
    nt: int = 64
    ds: float = 1.0
    dt: float = 0.0
    offset_s: float = 0.0
    offset_t: float = 0.0
    na: int = 64
    orbit: float = 360.0
    orbit_start: float = 0.0
    source_offset: float = 0.0
    dsd: float = 0.0
    dod: float = 0.0
    src: object = CtSourceCircle()

    is_arc = True

    def __post_init__(self):
        if not isinstance(self.src, (CtSourceCircle, CtSourceHelix)):
            raise TypeError("src must be a CtSourceCircle or a CtSourceHelix")
        if self.dt == 0.0:
            object.__setattr__(self, "dt", self.ds)
        if self.dsd == 0.0:
            object.__setattr__(self, "dsd", 4.0 * self.ns * self.ds)
        if self.dod == 0.0:
            object.__setattr__(self, "dod", 1.0 * self.ns * self.ds)

    def dims(self):
        return (self.ns, self.nt, self.na)

    @property
    def dso(self) -> float:
        return self.dsd - self.dod

    def angles(self):
        return self.orbit_start + np.arange(self.na) * (self.orbit / self.na)

    def ar(self, f32: bool = False):
        if f32:
            k = float(np.float32(math.pi / 180))
            deg = _f32(self.orbit_start) + np.arange(self.na) * (_f32(self.orbit) / self.na)
            return (deg * k).astype(np.float32).astype(np.float64)
        return self.angles() * (math.pi / 180)

    def ws(self) -> float:
        return (self.ns - 1) / 2 + _f32(self.offset_s)

    def wt(self) -> float:
        return (self.nt - 1) / 2 + _f32(self.offset_t)

    def s(self):
        """``_s(rg)`` (ct-geom.jl:22)."""
        return self.ds * (np.arange(self.ns) - self.ws())

    def t(self):
        """``_t(rg)`` (ct-geom.jl:23)."""
        return self.dt * (np.arange(self.nt) - self.wt())

    def gamma(self, s=None):
        s = self.s() if s is None else np.asarray(s, dtype=np.float64)
        return s / self.dsd if self.is_arc else np.arctan(s / self.dsd)

    def gamma_max(self):
        return float(np.max(np.abs(self.gamma())))

    def orbit_short(self):
        return 180 + 2 * math.degrees(self.gamma_max())

    def rfov(self):
        return self.dso * math.sin(self.gamma_max())

    @classmethod
    def short(cls, na: int = 128, orbit_start: float = 0.0, **kw):
        # type3.jl:333-341, 401-409
        tmp = cls(orbit=360.0, **kw)
        os_ = tmp.orbit_short()
        na2 = int(math.ceil(na * os_ / 360))
        return cls(na=na2, orbit=os_, orbit_start=orbit_start, **kw)

    def ones(self, dtype=np.float32):
        return np.ones(self.dims(), dtype=dtype, order="F")

    def cone_angle(self) -> float:
        """``_cone_angle`` = atan(nt dt / 2 / dsd) (ct-geom.jl:115)."""
        return math.atan(self.nt * self.dt / 2 / self.dsd)

    def zfov(self) -> float:
        """``_zfov`` = dso / dsd * nt * dt (ct-geom.jl:119)."""
        return self.dso / self.dsd * self.nt * self.dt

    def source_dz_per_view(self) -> float:
        """``_source_dz_per_view`` (ct-geom.jl:122-131): 0 for a circle; for a
        helix pitch * zfov / (na * 360 / orbit), and 0 for one view."""
        if not isinstance(self.src, CtSourceHelix) or self.na == 1:
            return 0.0
        na_per_360 = self.na * (360 / self.orbit)
        return self.src.pitch * self.zfov() / na_per_360

    def source_zs(self) -> np.ndarray:
        """``_source_zs`` (ct-geom.jl:134-140): the source z of every view."""
        if not isinstance(self.src, CtSourceHelix):
            return np.zeros(self.na)
        return np.arange(self.na) * self.source_dz_per_view() + self.src.source_z0


@dataclasses.dataclass(frozen=True)
class CtFanArc(CtFan):
    is_arc = True


@dataclasses.dataclass(frozen=True)
class CtFanFlat(CtFan):
    is_arc = False


def is_fan(rg) -> bool:
    return isinstance(rg, SinoFan)


def is_ct(rg) -> bool:
    return isinstance(rg, CtFan)


def is_full_orbit(orbit: float, period: float) -> bool:
    """``(orbit ÷ period) * period == orbit`` (parker.jl:35,112,148)."""
    return (orbit // period) * period == orbit


class ImageGeom:
    """Image grid: dims, deltas, offsets and a Boolean mask.

    ``axes()`` are the pixel centres ((0:n-1) - ((n-1)/2 + offset)) * delta,
    the ImageGeoms convention the reference passes into its backprojectors
    (``axes(ig)``, back-par.jl:46, cbct-back.jl:41).
    """

    def __init__(self, dims, deltas=None, offsets=None, mask=None):
        self.dims = tuple(int(n) for n in dims)
        nd = len(self.dims)
        self.deltas = tuple(float(d) for d in (deltas if deltas is not None else (1.0,) * nd))
        self.offsets = tuple(float(o) for o in (offsets if offsets is not None else (0.0,) * nd))
        if mask is None:
            mask = np.ones(self.dims, dtype=bool, order="F")
        mask = np.asfortranarray(np.asarray(mask, dtype=bool))
        if mask.shape != self.dims:
            raise ValueError("mask size mismatch")
        self.mask = mask

    def axes(self):
        return tuple(d * (np.arange(n) - ((n - 1) / 2 + o))
                     for n, d, o in zip(self.dims, self.deltas, self.offsets))

    def circle(self, r: float) -> np.ndarray:
        """Circular (cylindrical for 3-D) support mask of radius r (in the
        units of deltas), as ImageGeoms' ``circle(ig; r)``."""
        x, y = self.axes()[:2]
        m2 = (x[:, None] ** 2 + y[None, :] ** 2) <= r * r
        if len(self.dims) == 3:
            m2 = np.repeat(m2[:, :, None], self.dims[2], axis=2)
        return np.asfortranarray(m2)

    def with_mask(self, mask) -> "ImageGeom":
        return ImageGeom(self.dims, self.deltas, self.offsets, mask)

    def __repr__(self):
        return f"ImageGeom(dims={self.dims}, deltas={self.deltas}, offsets={self.offsets})"
