"""Test configuration: markers and import paths.

`gpu` tests need an MI355X and the built libsgx.so; the rest run on the CPU.
"""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "sinograms.jl_amd")
for p in (ROOT, PKG_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libsgx.so")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def sg():
    import sinograms_amd
    return sinograms_amd


def rel_rmse(x, ref):
    import numpy as np
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.sqrt(np.sum((x - ref) ** 2) / np.sum(ref ** 2)))


def c1_case(sg, short: bool):
    """The 2-D fan-beam workload C1: SinoFanArc, 888 bins of 1.0239 mm, dso 541,
    dsd 949; 984 views over 360 deg, or the short scan (642 views over 180 deg
    plus the fan angle, Parker weights); 512 x 512 at 0.98 mm inside the circle
    of radius rfov; 15 random ellipses (seed 1), 1e5-photon Poisson noise
    (seed 2, densities in units of 0.005 / mm, zero counts clamped to 1).
    Returns (rg, ig, sinogram Float32 (nb, na))."""
    import numpy as np
    mu = 0.005
    kw = dict(nb=888, d=1.0239, dsd=949.0, dod=949.0 - 541.0)
    full = sg.SinoFanArc(na=984, **kw)
    rg = sg.SinoFanArc(na=642, orbit=full.orbit_short(), **kw) if short else full
    ig = sg.ImageGeom((512, 512), (0.98, 0.98))
    ig = ig.with_mask(ig.circle(full.rfov()))
    p = sg.phantom.sino_ellipses(sg.phantom.random_ellipses(1, 15, 512 * 0.98), rg).numpy()
    noisy = (sg.phantom.poisson_log(mu * p, 1e5, 2) / mu).astype(np.float32)
    return rg, ig, noisy


def fan_references(orc, rg, ig, sino, Hk):
    """Three CPU evaluations of fbp(plan, sino) for an arc or flat fan
    (fbp.jl:32-38, 72-86): the Float32 restatement (the reference's
    precision), the Float64 one, and Float64 arithmetic on the Float32 path's
    rounded inputs (pixel centres x / ds, dsd / ds and dso / ds rounded to
    Float32, sin and cos of the Float32 angles rounded to Float32), which
    separates the Float32 path's arithmetic error from its input rounding."""
    import numpy as np
    f32 = np.float32
    ar = rg.ar(f32=True).astype(f32)
    vw = orc.angle_weight(ar)
    if rg.orbit < 360:
        vw = vw * orc.parker_fan_short(rg.gamma(), ar, rg.gamma_max())
    sw = orc.fan_sino_weight(rg.gamma(), rg.dso, rg.dsd)
    xc, yc = ig.axes()
    args = (Hk, vw, sw, ar, rg.dsd, rg.dso, 0.0, rg.is_arc, rg.d, 0.0, xc, yc, ig.mask)
    ref32 = orc.fbp_fan(sino, *args, dtype=f32)
    ref64 = orc.fbp_fan(sino.astype(np.float64), *args)
    s = sino.astype(np.float64) * np.asarray(sw, np.float64)[:, None] * np.asarray(vw, np.float64)
    q = orc.sino_filter(s, np.asarray(Hk, np.float64))
    a = ar.astype(np.float64)
    ds = f32(rg.d)
    ref64i = orc.back_fan(q, np.sin(a).astype(f32), np.cos(a).astype(f32),
                          float(f32(rg.dsd) / ds), float(f32(rg.dso) / ds), 0.0, rg.is_arc, 1.0,
                          0.0, (xc.astype(f32) / ds).astype(np.float64),
                          (yc.astype(f32) / ds).astype(np.float64), ig.mask)
    return ref32, ref64, ref64i
