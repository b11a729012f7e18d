"""sgx_cbct_back_noalloc (no device allocation, fixed device scratch) against
sgx_cbct_back, through the C ABI, on two calls in a row (the second smaller),
and its size and shape checks."""

import ctypes
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.dirname(HERE), os.path.join(os.path.dirname(HERE), "sinograms.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def _call(L, ctx, fn, rg, ig, proj, img, mask=None, ia_skip=1):
    betas = np.ascontiguousarray(rg.ar(f32=True), dtype=np.float64)
    xc, yc, zc = (np.ascontiguousarray(v, dtype=np.float64) for v in ig.axes())
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    return getattr(L, fn)(ctx.h, ctypes.c_void_p(proj.data_ptr()), rg.ns, rg.nt, rg.na, ia_skip,
                          P(betas), rg.dsd, rg.dso, rg.ds, rg.dt, rg.offset_s, rg.offset_t,
                          int(rg.is_arc), P(xc), len(xc), P(yc), len(yc), P(zc), len(zc),
                          None if mask is None else ctypes.c_void_p(mask.data_ptr()),
                          ctypes.c_void_p(img.data_ptr()), 0)


def test_noalloc_matches_cbct_back():
    import torch
    import sinograms_amd as sg
    assert torch.cuda.is_available()
    L = sg._lib.load()
    ctx = sg._lib.context(0)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(77)
    cases = [
        # the FDK workload's geometry at full size (a 256^2 x 128 volume, 360 views)
        (sg.CtFanFlat(ns=512, nt=384, ds=1.2, dt=1.2, na=360, dsd=1200.0, dod=400.0),
         sg.ImageGeom((256, 256, 128), (1.0, 1.0, 1.0)), 1, False),
        # then a smaller arc-detector call with a mask and view skipping
        (sg.CtFanArc(ns=64, nt=40, ds=1.0, dt=1.2, na=32, offset_s=0.25),
         sg.ImageGeom((30, 28, 26), (1.0, 1.0, 1.0)), 2, True),
    ]
    for rg, ig, skip, masked in cases:
        print("noalloc", type(rg).__name__, rg.dims(), ig.dims, skip, masked, flush=True)
        proj = torch.from_numpy(rng.standard_normal(rg.dims()[::-1]).astype(np.float32)).to(dev)
        proj = proj.permute(2, 1, 0)                      # (ns, nt, na) column-major
        mask = None
        if masked:
            m = rng.random(ig.dims) > 0.3
            mask = torch.from_numpy(np.ascontiguousarray(m.astype(np.uint8).T)).to(dev)
        want = torch.full(ig.dims[::-1], np.nan, device=dev)
        got = torch.full(ig.dims[::-1], np.nan, device=dev)
        ctx.check(_call(L, ctx, "sgx_cbct_back", rg, ig, proj, want, mask, skip))
        ctx.check(_call(L, ctx, "sgx_cbct_back_noalloc", rg, ig, proj, got, mask, skip))
        assert torch.isfinite(want).all()
        assert torch.equal(got, want)


def test_noalloc_rejects_what_it_cannot_take():
    import torch
    import sinograms_amd as sg
    L = sg._lib.load()
    ctx = sg._lib.context(0)
    dev = torch.device("cuda:0")
    rg = sg.CtFanFlat(ns=64, nt=40, ds=1.0, dt=1.2, na=360)
    proj = torch.zeros(rg.dims()[::-1], device=dev).permute(2, 1, 0)
    # a window table beyond the scratch: 512 x 512 x 256 voxels x 360 views
    ig = sg.ImageGeom((512, 512, 256), (1.0, 1.0, 1.0))
    img = torch.zeros(ig.dims[::-1], device=dev)
    rc = _call(L, ctx, "sgx_cbct_back_noalloc", rg, ig, proj, img)
    assert rc == sg._lib.SGX_ERR_UNSUPPORTED
    assert b"scratch" in L.sgx_last_error(ctx.h)
    # a detector row that is not a multiple of 4 bins
    rg2 = sg.CtFanFlat(ns=63, nt=40, ds=1.0, dt=1.2, na=8)
    p2 = torch.zeros(rg2.dims()[::-1], device=dev).permute(2, 1, 0)
    ig2 = sg.ImageGeom((8, 8, 8), (1.0, 1.0, 1.0))
    rc = _call(L, ctx, "sgx_cbct_back_noalloc", rg2, ig2, p2, torch.zeros(8, 8, 8, device=dev))
    assert rc == sg._lib.SGX_ERR_UNSUPPORTED
