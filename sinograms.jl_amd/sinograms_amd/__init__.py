"""sinograms_amd: the Sinograms.jl reconstruction hot path on MI355X.

Host-side mirror of the reference's API (geometries, plan_fbp, fbp, fdk,
fbp_back, cbct_back, fbp_sino_filter, parker_weight, ...) over libsgx.so,
the HIP library whose C ABI is include/sgx.h.  The device work always runs
in libsgx; there is no CPU fallback.
"""

from .geom import (CtFan, CtFanArc, CtFanFlat, CtSourceCircle, CtSourceHelix, ImageGeom,
                   SinoFan, SinoFanArc, SinoFanFlat, SinoGeom, SinoMoj, SinoPar)
from .recon import (Boxcar, FBPNormalPlan, FDKplan, MojPlan, Hamming, Hann, NoWindow, SizeMismatch,
                    Window, WindowVect, _angle_weights, _fbp_weights, _fdk_weights, cbct_back,
                    fbp, fbp_back, fbp_filter, fbp_ramp, fbp_sino_filter, fbp_sino_weight,
                    fbp_window, fdk, fdk_weight_cyl, parker_weight, parker_weight_fan_short,
                    parker_weight_par, plan_fbp, ramp_arc, ramp_flat)
from . import phantom
from .batch import Batch, autotune, tuning
from .sys2 import backproject_bdd, project_bdd
from . import raygeom
from .raygeom import footprint_size, rays
from ._lib import LIB_PATH, SgxError

__all__ = [
    "SinoGeom", "SinoPar", "SinoMoj", "MojPlan", "SinoFan", "SinoFanArc", "SinoFanFlat", "CtFan", "CtFanArc",
    "CtFanFlat", "CtSourceCircle", "CtSourceHelix", "ImageGeom", "plan_fbp", "fbp", "fdk", "fbp_back", "cbct_back",
    "fbp_sino_filter", "fbp_filter", "fbp_ramp", "ramp_flat", "ramp_arc", "fbp_window",
    "Window", "NoWindow", "Boxcar", "Hamming", "Hann", "WindowVect", "parker_weight",
    "parker_weight_par", "parker_weight_fan_short", "fbp_sino_weight", "fdk_weight_cyl",
    "FBPNormalPlan", "FDKplan", "SizeMismatch", "SgxError", "phantom", "project_bdd",
    "backproject_bdd", "Batch", "autotune", "tuning", "rays", "footprint_size", "raygeom",
]
