"""GPU parity tests: libsgx (through its C ABI) against the CPU oracle.

Bars: rel-RMSE <= 1e-5 against the oracle's reference-precision (Float32)
mode for the FP32 paths; the BF16-storage form is checked exactly against
the oracle run on the same bf16-rounded data (so only accumulation order
differs) and its storage error against Float64 is reported.
"""

import ctypes
import math
import os
import sys

import numpy as np
import pytest

from conftest import rel_rmse

pytestmark = pytest.mark.gpu

TOL32 = 1e-5


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def _grid(n, d=1.0, off=0.0):
    return d * (np.arange(n) - ((n - 1) / 2 + off))


def _bf16_round(x):
    u = np.ascontiguousarray(np.asarray(x, np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(np.shape(x))


# --- K3 cone-beam backprojection -------------------------------------------------------

@pytest.fixture(params=["staged", "gather", "staged_z8", "staged_z8w"])
def cbct_kernel(request, sg, torch_cuda):
    """Run a test with each cone-beam backprojector of the shared context."""
    ctx = sg._lib.context(0)
    old = ctx.cbct_kernel
    ctx.cbct_kernel = request.param
    yield request.param
    ctx.cbct_kernel = old


@pytest.mark.parametrize("arc", [False, True])
@pytest.mark.parametrize("dims,ia_skip,masked", [((16, 16, 8), 1, False),
                                                 ((19, 13, 11), 2, True),
                                                 ((9, 24, 35), 1, True),
                                                 ((37, 21, 40), 1, False)])
def test_cbct_back_fp32_vs_oracle(sg, orc, torch_cuda, cbct_kernel, arc, dims, ia_skip, masked):
    _say("cbct_back", cbct_kernel, arc, dims, ia_skip, masked)
    rng = np.random.default_rng(10)
    ns, nt, na = 40, 30, 24
    proj = rng.standard_normal((ns, nt, na)).astype(np.float32)
    rg = (sg.CtFanArc if arc else sg.CtFanFlat)(ns=ns, nt=nt, ds=1.1, dt=1.3, na=na,
                                                 offset_s=0.25, offset_t=-0.5, dsd=120.0,
                                                 dod=50.0)
    ig = sg.ImageGeom(dims, (1.0, 0.9, 1.1), (0.3, 0.0, -0.2))
    if masked:
        m = rng.random(dims) > 0.3
        ig = ig.with_mask(m)
    got = sg.cbct_back(proj, rg, ig, ia_skip=ia_skip)
    xc, yc, zc = ig.axes()
    ref = orc.cbct_back(proj, rg.ar(f32=True), rg.dsd, rg.dso, rg.ds, rg.dt, rg.offset_s,
                        rg.offset_t, arc, xc, yc, zc, ig.mask, ia_skip=ia_skip,
                        dtype=np.float32)
    assert got.shape == dims
    assert np.all(got[~ig.mask] == 0)
    assert rel_rmse(got, ref) < TOL32
    # device pointers give the same answer as host pointers
    t = torch_cuda.from_numpy(np.ascontiguousarray(proj.transpose(2, 1, 0))).cuda().permute(2, 1, 0)
    got_d = sg.cbct_back(t, rg, ig, ia_skip=ia_skip).cpu().numpy()
    assert np.array_equal(got_d, got)


@pytest.mark.parametrize("arc", [False, True])
@pytest.mark.parametrize("case", ["interior", "edges", "masked_skip", "exact_t"])
def test_cbct_back_mirror_vs_oracle(sg, orc, torch_cuda, arc, case):
    # The mirror form of the staged kernel: with the z axis symmetric about 0,
    # offset_t = 0 and no source z, one detector position serves a voxel
    # z > 0 and its image at -z (bt(-z) = nt + 1 - bt(z)).  "edges": a short
    # detector, so that windows at the top and bottom reach its edge and
    # lanes take the gather path; "exact_t": the exact arc t magnification,
    # which keeps the symmetry.  Against the Float32 oracle, and against the
    # same kernel with the mirror form off.
    _say("cbct_back mirror", arc, case)
    rng = np.random.default_rng(50)
    nt = 20 if case == "edges" else 48
    rg = (sg.CtFanArc if arc else sg.CtFanFlat)(ns=64, nt=nt, ds=1.1, dt=1.3, na=36,
                                                 offset_s=0.25, offset_t=0.0, dsd=120.0,
                                                 dod=50.0)
    proj = rng.standard_normal(rg.dims()).astype(np.float32)
    dims = (24, 20, 32)
    ig = sg.ImageGeom(dims, (1.0, 0.9, 1.1), (0.3, 0.0, 0.0))
    ia_skip = 1
    if case == "masked_skip":
        ig = ig.with_mask(rng.random(dims) > 0.3)
        ia_skip = 2
    exact = case == "exact_t"
    xc, yc, zc = ig.axes()
    assert np.array_equal(zc[::-1], -zc)
    ctx = sg._lib.context(0)
    try:
        ctx.cbct_mirror = "always"
        got = sg.cbct_back(proj, rg, ig, ia_skip=ia_skip, arc_t_exact=exact)
        ctx.cbct_mirror = "never"
        got0 = sg.cbct_back(proj, rg, ig, ia_skip=ia_skip, arc_t_exact=exact)
    finally:
        ctx.cbct_mirror = "fp16"
    ref = orc.cbct_back(proj, rg.ar(f32=True), rg.dsd, rg.dso, rg.ds, rg.dt, rg.offset_s,
                        rg.offset_t, arc, xc, yc, zc, ig.mask, ia_skip=ia_skip,
                        dtype=np.float32, arc_t_exact=exact)
    e, e0 = rel_rmse(got, ref), rel_rmse(got, got0)
    _say(f"  vs Float32 oracle {e:.3g}, vs the form without mirror {e0:.3g}")
    assert np.all(got[~ig.mask] == 0)
    assert e < TOL32
    assert e0 < 2e-6
    if case == "edges":
        assert np.any(ref == 0)          # voxels that see no detector row in some view


@pytest.mark.parametrize("prec", ["fp32", "f16s"])
def test_fdk_view_chunks(sg, orc, torch_cuda, prec):
    # A cone-beam execute in view chunks (the filters of later chunks on the
    # plan's side stream, beside the backprojections of earlier chunks)
    # against one filter launch and one backprojection: per-chunk sums added
    # in chunk order, so equal to FP32 rounding; repeated device executes,
    # one with its input produced on another stream, equal the first bit for
    # bit; and the chunked result stays within the FP32 bar of the oracle.
    import torch
    _say("fdk view chunks", prec)
    rg = sg.CtFanFlat(ns=256, nt=48, ds=1.2, dt=1.2, na=100, dsd=600.0, dod=200.0)
    ig = sg.ImageGeom((48, 44, 32), (1.0, 1.0, 1.0))
    ell = sg.phantom.random_ellipsoids(61, 8, 48.0)
    proj = sg.phantom.proj_ellipsoids(ell, rg).numpy().astype(np.float32)
    dev = _colmajor(torch_cuda, proj)
    ctx = sg._lib.context(0)
    vols, plans = {}, {}
    before = ctx.view_chunks
    try:
        for k in (1, 2, 3):
            ctx.view_chunks = k
            plans[k] = sg.plan_fbp(rg, ig, precision=prec)
            vols[k] = sg.fdk(plans[k], proj)
    finally:
        ctx.view_chunks = before
    for k in (2, 3):
        e = rel_rmse(vols[k], vols[1])
        _say(f"  {k} chunks vs 1: rel-RMSE {e:.3g}")
        assert e < 1e-6, (k, e)
    plan = plans[3]
    outs = [plan.execute(dev) for _ in range(3)]
    for o in outs:
        assert np.array_equal(o.cpu().numpy(), vols[3])
    side = torch.cuda.Stream()
    late = torch.zeros_like(dev)
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        torch.cuda._sleep(20_000_000)          # keeps `side` busy for several ms
        late.copy_(dev)
        got = plan.execute(late).cpu().numpy()
    assert np.array_equal(got, vols[3])
    if prec == "fp32":
        ar = rg.ar(f32=True).astype(np.float32)
        vw = orc.fdk_weight_cyl(rg.s(), rg.t(), rg.dsd, rg.dso, False)[:, :, None] * \
            orc.angle_weight(ar)
        xc, yc, zc = ig.axes()
        ref = orc.fdk(proj, plan.filter, vw, ar, rg.dsd, rg.dso, rg.ds, rg.dt, 0.0, 0.0, False,
                      xc, yc, zc, ig.mask, dtype=np.float32)
        e = rel_rmse(vols[3], ref)
        _say(f"  3 chunks vs Float32 oracle: rel-RMSE {e:.3g}")
        assert e < TOL32


@pytest.mark.parametrize("prec", ["fp32", "bf16", "f16s"])
@pytest.mark.parametrize("voxels", ["fine", "coarse"])
def test_fdk_block_forms_bit_identical(sg, orc, torch_cuda, prec, voxels):
    # The staged kernel's 8-deep blocks (with the usual or the wide windows)
    # run every voxel through the same arithmetic as the 16-deep blocks.
    # "fine": every view of every block is staged, views summed in order, and
    # the volumes are bit for bit the same (nz = 40: a partial block at the
    # top).  "coarse": 1.8 mm voxels at magnification 2 on 1 mm bins; a
    # 16 x 16 column block spans about 55 bins, past the 48-bin windows (those
    # lanes gather the views they miss after the staged ones, so the sums'
    # order differs) and inside the wide windows, which stage every view.
    _say("fdk block forms", prec, voxels)
    rng = np.random.default_rng(51)
    if voxels == "fine":
        rg = sg.CtFanFlat(ns=96, nt=64, ds=1.2, dt=1.2, na=48, dsd=400.0, dod=150.0)
        ig = sg.ImageGeom((40, 36, 40), (1.0, 1.0, 1.0))
    else:
        rg = sg.CtFanFlat(ns=160, nt=96, ds=1.0, dt=1.0, na=36, dsd=400.0, dod=200.0)
        ig = sg.ImageGeom((40, 36, 24), (1.8, 1.8, 1.8))
    if voxels == "fine":
        proj = rng.standard_normal(rg.dims()).astype(np.float32)
    else:                       # a phantom: the oracle check below at the FP32 bar
        ell = sg.phantom.random_ellipsoids(51, 8, 60.0)
        proj = sg.phantom.proj_ellipsoids(ell, rg).numpy().astype(np.float32)
    ctx = sg._lib.context(0)
    vols = {}
    try:
        ctx.cbct_mirror = "never"
        for k in ("staged", "staged_z8", "staged_z8w"):
            ctx.cbct_kernel = k
            plan = sg.plan_fbp(rg, ig, precision=prec)
            assert sg.tuning(plan)["cbct_kernel"] == k
            vols[k] = sg.fdk(plan, proj)
    finally:
        ctx.cbct_kernel = "staged"
        ctx.cbct_mirror = "fp16"
    if voxels == "fine":
        assert np.array_equal(vols["staged"], vols["staged_z8"])
        assert np.array_equal(vols["staged"], vols["staged_z8w"])
        return
    for k in ("staged_z8", "staged_z8w"):
        assert rel_rmse(vols[k], vols["staged"]) < 1e-6
    if prec == "fp32":
        ar = rg.ar(f32=True).astype(np.float32)
        vw = orc.fdk_weight_cyl(rg.s(), rg.t(), rg.dsd, rg.dso, False)[:, :, None] * \
            orc.angle_weight(ar)
        xc, yc, zc = ig.axes()
        ref = orc.fdk(proj, plan.filter, vw, ar, rg.dsd, rg.dso, rg.ds, rg.dt, 0.0, 0.0, False,
                      xc, yc, zc, ig.mask, dtype=np.float32)
        assert rel_rmse(vols["staged_z8w"], ref) < TOL32


def test_cbct_back_detector_edges(sg, orc, cbct_kernel):
    # voxels projecting beyond every detector edge, including NaN-free far field
    _say("cbct_back edges", cbct_kernel)
    rng = np.random.default_rng(11)
    rg = sg.CtFanFlat(ns=12, nt=10, ds=2.0, na=8, dsd=60.0, dod=20.0)
    proj = rng.standard_normal(rg.dims()).astype(np.float32)
    ig = sg.ImageGeom((32, 30, 10), (1.0, 1.0, 2.0))
    got = sg.cbct_back(proj, rg, ig)
    xc, yc, zc = ig.axes()
    ref = orc.cbct_back(proj, rg.ar(f32=True), rg.dsd, rg.dso, rg.ds, rg.dt, 0.0, 0.0, False,
                        xc, yc, zc, ig.mask, dtype=np.float32)
    assert np.isfinite(got).all()
    assert rel_rmse(got, ref) < TOL32
    assert np.any(ref == 0)          # some voxels see no detector at all


@pytest.mark.parametrize("arc", [False, True])
@pytest.mark.parametrize("helix,exact", [(True, False), (False, True), (True, True)])
def test_cbct_back_options_vs_oracle(sg, orc, torch_cuda, cbct_kernel, arc, helix, exact):
    # the source z of every view (a helix) and the exact arc-detector t
    # magnification, which the reference keeps commented out
    # (cbct-back.jl:36-59, 180, 188); many voxels leave the detector in t
    _say("cbct_back options", cbct_kernel, arc, helix, exact)
    rng = np.random.default_rng(12)
    src = sg.CtSourceHelix(pitch=0.8, source_z0=-6.0) if helix else sg.CtSourceCircle()
    rg = (sg.CtFanArc if arc else sg.CtFanFlat)(ns=48, nt=36, ds=1.1, dt=1.3, na=30,
                                                 offset_s=0.25, offset_t=-0.5, dsd=120.0,
                                                 dod=50.0, src=src)
    proj = rng.standard_normal(rg.dims()).astype(np.float32)
    ig = sg.ImageGeom((29, 23, 26), (1.0, 0.9, 1.1), (0.3, 0.0, -0.2))
    xc, yc, zc = ig.axes()
    zs = rg.source_zs() if helix else None
    for ia_skip in (1, 2):
        got = sg.cbct_back(proj, rg, ig, ia_skip=ia_skip, use_source_zs=helix,
                           arc_t_exact=exact)
        ref = orc.cbct_back(proj, rg.ar(f32=True), rg.dsd, rg.dso, rg.ds, rg.dt, rg.offset_s,
                            rg.offset_t, arc, xc, yc, zc, ig.mask, ia_skip=ia_skip,
                            dtype=np.float32, source_zs=zs, arc_t_exact=exact)
        assert rel_rmse(got, ref) < TOL32, ia_skip
    base = sg.cbct_back(proj, rg, ig, ia_skip=2)
    if helix or arc:
        assert rel_rmse(base, got) > 1e-3           # the options change the geometry
    else:
        assert np.array_equal(base, got)             # arc_t_exact leaves a flat detector


@pytest.mark.parametrize("prec", ["fp32", "f16s"])
def test_fdk_plan_options_vs_oracle(sg, orc, torch_cuda, prec):
    _say("fdk plan options", prec)
    rng = np.random.default_rng(13)
    rg = sg.CtFanArc(ns=64, nt=40, ds=1.0, dt=1.2, na=32,
                     src=sg.CtSourceHelix(pitch=0.5, source_z0=-3.0))
    ig = sg.ImageGeom((30, 28, 26), (1.0, 1.0, 1.0))
    proj = rng.standard_normal(rg.dims()).astype(np.float32)
    opts = dict(use_source_zs=True, arc_t_exact=True)
    plan = sg.plan_fbp(rg, ig, precision=prec, **opts)
    got = sg.fdk(plan, proj)
    ar = rg.ar(f32=True).astype(np.float32)
    vw = orc.fdk_weight_cyl(rg.s(), rg.t(), rg.dsd, rg.dso, True)[:, :, None] * \
        orc.angle_weight(ar)
    xc, yc, zc = ig.axes()
    args = (plan.filter, vw, ar, rg.dsd, rg.dso, rg.ds, rg.dt, 0.0, 0.0, True, xc, yc, zc,
            ig.mask)
    back = dict(source_zs=rg.source_zs(), arc_t_exact=True)
    if prec == "fp32":
        e32 = rel_rmse(got, orc.fdk(proj, *args, dtype=np.float32, **back))
        _say(f"  rel-RMSE vs Float32 oracle {e32:.3g}")
        assert e32 < TOL32
    else:
        e = rel_rmse(got, sg.fdk(sg.plan_fbp(rg, ig, **opts), proj))
        _say(f"  F16S vs FP32 with the options {e:.3g}")
        assert e < 1e-3
    # the plan's default is the reference's geometry
    assert rel_rmse(sg.fdk(sg.plan_fbp(rg, ig, precision=prec), proj), got) > 1e-3


@pytest.mark.parametrize("arc", [False, True])
def test_cbct_back_staged_window_overflow(sg, orc, cbct_kernel, arc):
    # Fine detector pitch: a 16 x 16 x 16 block's footprint (~75 bins x ~50
    # rows) exceeds the staged window, so some lanes read from the window and
    # the rest gather from global memory within the same view.
    _say("cbct_back window overflow", cbct_kernel, arc)
    rng = np.random.default_rng(13)
    rg = (sg.CtFanArc if arc else sg.CtFanFlat)(ns=200, nt=150, ds=0.45, dt=0.45, na=30,
                                                 dsd=300.0, dod=100.0)
    ig = sg.ImageGeom((40, 40, 20), (1.0, 1.0, 1.0))
    proj = rng.standard_normal((200, 150, 30)).astype(np.float32)
    got = sg.cbct_back(proj, rg, ig)
    xc, yc, zc = ig.axes()
    ref = orc.cbct_back(proj, rg.ar(f32=True), rg.dsd, rg.dso, rg.ds, rg.dt, 0.0, 0.0, arc,
                        xc, yc, zc, ig.mask, dtype=np.float32)
    assert rel_rmse(got, ref) < TOL32


def test_cbct_back_bf16_storage(sg, orc, torch_cuda):
    _say("cbct_back bf16")
    torch = torch_cuda
    rng = np.random.default_rng(12)
    rg = sg.CtFanArc(ns=48, nt=32, ds=1.0, dt=1.2, na=30, dsd=150.0, dod=60.0)
    proj = rng.standard_normal(rg.dims()).astype(np.float32)
    pb = _bf16_round(proj)
    tb = torch.from_numpy(np.ascontiguousarray(proj.transpose(2, 1, 0))).cuda().to(
        torch.bfloat16)
    ig = sg.ImageGeom((20, 18, 12))
    from sinograms_amd import _lib
    import ctypes
    L = _lib.load()
    ctx = _lib.context(0)
    out = torch.empty((12, 18, 20), device="cuda")
    xc, yc, zc = (np.ascontiguousarray(v) for v in ig.axes())
    betas = np.ascontiguousarray(rg.ar(f32=True))
    p = lambda a: ctypes.c_void_p(a.ctypes.data)
    ctx.check(L.sgx_cbct_back(ctx.h, ctypes.c_void_p(tb.data_ptr()), rg.ns, rg.nt, rg.na, 1,
                              p(betas), rg.dsd, rg.dso, rg.ds, rg.dt, 0.0, 0.0, 1, p(xc), 20,
                              p(yc), 18, p(zc), 12, None, ctypes.c_void_p(out.data_ptr()),
                              _lib.SGX_Q_BF16))
    got = out.permute(2, 1, 0).cpu().numpy()
    ref = orc.cbct_back(pb, betas, rg.dsd, rg.dso, 1.0, 1.2, 0.0, 0.0, True, xc, yc, zc,
                        ig.mask, dtype=np.float32)
    assert rel_rmse(got, ref) < TOL32


# --- K1 / K2 2-D backprojection ----------------------------------------------------------

@pytest.mark.parametrize("kind", ["par", "arc", "flat"])
@pytest.mark.parametrize("nx,ny,na,ia_skip", [(16, 12, 30, 1), (37, 41, 50, 3),
                                              (128, 128, 180, 1)])
def test_back2d_fp32_vs_oracle(sg, orc, torch_cuda, kind, nx, ny, na, ia_skip):
    _say("back2d", kind, nx, ny, na, ia_skip)
    rng = np.random.default_rng(20)
    nb = 61
    if kind == "par":
        rg = sg.SinoPar(nb=nb, d=1.3, na=na, offset=0.25)
    else:
        rg = (sg.SinoFanArc if kind == "arc" else sg.SinoFanFlat)(
            nb=nb, d=1.3, na=na, offset=0.25, dsd=300.0, dod=100.0, source_offset=0.7)
    sino = rng.standard_normal(rg.dims()).astype(np.float32)
    ig = sg.ImageGeom((nx, ny), (0.8, 0.9))
    ig = ig.with_mask(rng.random((nx, ny)) > 0.2)
    got = sg.fbp_back(rg, ig, sino, ia_skip=ia_skip)
    xc, yc = ig.axes()
    ar32 = rg.ar(f32=True)[::ia_skip].astype(np.float32)
    sn, cs = np.sin(ar32), np.cos(ar32)
    if kind == "par":
        ref = orc.back_par(sino, sn, cs, rg.d, rg.offset, xc, yc, ig.mask, ia_skip=ia_skip,
                           dtype=np.float32)
    else:
        ref = orc.back_fan(sino, sn, cs, rg.dsd, rg.dso, rg.source_offset, kind == "arc",
                           rg.d, rg.offset, xc, yc, ig.mask, ia_skip=ia_skip,
                           dtype=np.float32)
    assert np.all(got[~ig.mask] == 0)
    assert rel_rmse(got, ref) < TOL32


@pytest.mark.parametrize("kind", ["par", "arc", "flat"])
def test_back2d_window_fallbacks(sg, orc, torch_cuda, kind):
    # The windowed kernel's two fallbacks to global reads: a 16-pixel tile
    # that spans more bins than a window slot holds (fine detector pitch),
    # and, for the fan beams, pixels at or behind the source (dso smaller
    # than the image), where the tile's corner bound does not hold.
    _say("back2d window fallbacks", kind)
    rng = np.random.default_rng(21)
    if kind == "par":
        rg = sg.SinoPar(nb=1500, d=0.1, na=40, offset=0.5)
        ig = sg.ImageGeom((96, 80), (1.0, 1.1))
    else:
        rg = (sg.SinoFanArc if kind == "arc" else sg.SinoFanFlat)(
            nb=700, d=0.25, na=48, dsd=60.0, dod=20.0, offset=0.5)
        ig = sg.ImageGeom((112, 96), (1.0, 1.0))
        # leave out a ring around the source orbit, where w2 = dsd^2 / (d^2 + r^2)
        # amplifies FP32 rounding of d and r without bound
        xc, yc = ig.axes()
        rad = np.hypot(xc[:, None], yc[None, :])
        ig = ig.with_mask(np.abs(rad - rg.dso) > 3.0)
    sino = rng.standard_normal(rg.dims()).astype(np.float32)
    got = sg.fbp_back(rg, ig, sino)
    xc, yc = ig.axes()
    ar32 = rg.ar(f32=True).astype(np.float32)
    sn, cs = np.sin(ar32), np.cos(ar32)
    refs = {}
    for dt in (np.float32, np.float64):
        q, a, b = sino.astype(dt), sn.astype(dt), cs.astype(dt)
        if kind == "par":
            refs[dt] = orc.back_par(q, a, b, rg.d, rg.offset, xc, yc, ig.mask, dtype=dt)
        else:
            refs[dt] = orc.back_fan(q, a, b, rg.dsd, rg.dso, 0.0, kind == "arc", rg.d,
                                    rg.offset, xc, yc, ig.mask, dtype=dt)
    # random data on a fine detector: the reference-precision (Float32) path
    # is itself far from Float64 here, so the bar is relative to its error
    floor = rel_rmse(refs[np.float32], refs[np.float64])
    e64 = rel_rmse(got, refs[np.float64])
    _say(f"  rel-RMSE vs Float64 {e64:.3g} (Float32 oracle {floor:.3g})")
    assert np.abs(refs[np.float64]).max() > 0
    assert e64 < 2 * floor + 1e-6
    assert rel_rmse(got, refs[np.float32]) < max(TOL32, 3 * floor)


# --- K4 filter and K5 Parker ------------------------------------------------------------

@pytest.mark.parametrize("nb,ncols,npad,extra", [(37, 5, 64, 3), (367, 360, 512, 0),
                                                 (888, 70, 1024, 0), (512, 1000, 1024, 2),
                                                 (300, 37, 1024, 3), (512, 33, 1024, 0),
                                                 (1024, 9, 1024, 0), (1500, 40, 2048, 0),
                                                 (1024, 33, 2048, 0), (2047, 9, 2048, 1),
                                                 (1, 3, 1024, 0),
                                                 (3000, 40, 4096, 5), (8192, 6, 8192, 0)])
def test_sino_filter_fp32_vs_oracle(sg, orc, torch_cuda, nb, ncols, npad, extra):
    _say("filter fp32", nb, ncols, npad, extra)
    rng = np.random.default_rng(30)
    x = rng.standard_normal((nb, ncols)).astype(np.float32)
    h, _ = orc.ramp_arc(npad, 1.0, max(900.0, float(npad)))
    Hk = orc.fbp_filter(h, npad, 1.0, window=orc.fbp_window("hann", npad)).astype(np.float32)
    got = sg.fbp_sino_filter(x, Hk, extra=extra)
    ref = orc.sino_filter(x, Hk, extra=extra, dtype=np.float32)
    assert got.shape == (nb + extra, ncols)
    assert np.all(got[nb:] == 0)
    assert rel_rmse(got, ref) < 2e-6


@pytest.mark.parametrize("nb,ncols,npad", [(37, 5, 64), (367, 360, 512), (512, 1000, 1024),
                                          (3000, 40, 4096)])
def test_sino_filter_split_bf16x3(sg, orc, torch_cuda, nb, ncols, npad):
    _say("filter bf16x3", nb, ncols, npad)
    rng = np.random.default_rng(31)
    x = rng.standard_normal((nb, ncols)).astype(np.float32)
    h, _ = orc.ramp_flat(npad, 1.0)
    Hk = orc.fbp_filter(h, npad, 1.0).astype(np.float32)
    got = sg.fbp_sino_filter(x, Hk, precision="bf16x3")
    ref = orc.sino_filter(x.astype(np.float64), Hk.astype(np.float64))
    err = rel_rmse(got, ref)
    _say("  split-bf16x3 rel-RMSE vs Float64:", err)
    assert err < 3e-5


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_sino_filter_wide_many_columns(sg, orc, torch_cuda, precision):
    # 2500 bins x 33000 columns: the wide-detector instance with 8 row tiles
    # per wave (more than 1024 column blocks); the columns are independent,
    # so a sample of them is checked against the oracle.
    _say("filter wide, many columns", precision)
    nb, ncols, npad = 2500, 33000, 4096
    rng = np.random.default_rng(32)
    x = rng.standard_normal((nb, ncols)).astype(np.float32)
    h, _ = orc.ramp_arc(npad, 1.0, 1500.0)
    Hk = orc.fbp_filter(h, npad, 1.0).astype(np.float32)
    got = sg.fbp_sino_filter(x, Hk, extra=3, precision=precision)
    assert got.shape == (nb + 3, ncols)
    cols = np.unique(np.r_[0, 1, 31, 32, ncols - 1, rng.integers(0, ncols, 40)])
    ref = orc.sino_filter(x[:, cols].astype(np.float64), Hk.astype(np.float64), extra=3)
    err = rel_rmse(got[:, cols], ref)
    _say(f"  rel-RMSE vs Float64 over {len(cols)} columns: {err:.3g}")
    assert np.all(got[nb:] == 0)
    assert err < (2e-6 if precision == "fp32" else 3e-5)


def test_wide_detector_plans_vs_oracle(sg, orc, torch_cuda):
    # Fan and cone plans whose detectors are wider than 2048 bins.  With
    # thousands of fine bins the Float32 bin coordinate itself limits FP32
    # accuracy (the Float32 oracle is ~1e-4 from Float64 on a noise
    # sinogram here), so the bar is: as close to the Float64 oracle as the
    # Float32 oracle is, within a factor 1.5.
    _say("wide-detector fan and cone plans")
    rng = np.random.default_rng(33)
    rg = sg.SinoFanArc(nb=2600, na=24, d=0.25, dsd=1100.0, dod=400.0)
    ig = sg.ImageGeom((48, 40), (4.0, 4.0))
    sino = rng.standard_normal(rg.dims()).astype(np.float32)
    plan = sg.plan_fbp(rg, ig)
    got = sg.fbp(plan, sino)
    ar = rg.ar(f32=True).astype(np.float32)
    vw = orc.angle_weight(ar)
    sw = orc.fan_sino_weight(rg.gamma(), rg.dso, rg.dsd)
    xc, yc = ig.axes()
    args = (plan.filter, vw, sw, ar, rg.dsd, rg.dso, 0.0, True, rg.d, 0.0, xc, yc, ig.mask)
    ref64 = orc.fbp_fan(sino.astype(np.float64), *args)
    e, e_orc = rel_rmse(got, ref64), rel_rmse(orc.fbp_fan(sino, *args, dtype=np.float32), ref64)
    _say(f"  fan nb=2600 rel-RMSE vs Float64 oracle {e:.3g} (Float32 oracle {e_orc:.3g})")
    assert e < 1.5 * e_orc + 1e-6

    cg = sg.CtFanFlat(ns=2304, nt=6, ds=0.2, dt=0.2, na=12, dsd=1200.0, dod=400.0)
    cig = sg.ImageGeom((40, 36, 4), (4.0, 4.0, 0.15))
    proj = rng.standard_normal(cg.dims()).astype(np.float32)
    p32 = sg.plan_fbp(cg, cig)
    r32 = sg.fdk(p32, proj)
    car = cg.ar(f32=True).astype(np.float32)
    cvw = orc.fdk_weight_cyl(cg.s(), cg.t(), cg.dsd, cg.dso, False)[:, :, None] * \
        orc.angle_weight(car)
    cx, cy, cz = cig.axes()
    cargs = (p32.filter, cvw, car, cg.dsd, cg.dso, cg.ds, cg.dt, 0.0, 0.0, False, cx, cy, cz,
             cig.mask)
    cref64 = orc.fdk(proj.astype(np.float64), *cargs)
    e = rel_rmse(r32, cref64)
    e_orc = rel_rmse(orc.fdk(proj, *cargs, dtype=np.float32), cref64)
    _say(f"  cone ns=2304 FP32 rel-RMSE vs Float64 oracle {e:.3g} (Float32 oracle {e_orc:.3g})")
    assert e < 1.5 * e_orc + 1e-6
    for prec, tol in (("f16s", 1e-3), ("bf16", 1e-2)):
        e = rel_rmse(sg.fdk(sg.plan_fbp(cg, cig, precision=prec), proj), r32)
        _say(f"  cone ns=2304 {prec} vs FP32 {e:.3g}")
        assert e < tol


def _toeplitz_plan(sg, rg, ig, **kw):
    """The same plan with the Toeplitz (MFMA) form of the filter."""
    os.environ["SGX_FILTER_TOEPLITZ"] = "1"
    try:
        return sg.plan_fbp(rg, ig, **kw)
    finally:
        del os.environ["SGX_FILTER_TOEPLITZ"]


@pytest.mark.parametrize("arc", [True, False])
@pytest.mark.parametrize("short", [False, True])
@pytest.mark.parametrize("nb", [400, 600, 1300])
def test_fft_filter_fan_plan_vs_oracle(sg, orc, torch_cuda, arc, short, nb):
    # 400 / 600 / 1300 bins: npad = nextpow2(nb + 1) = 512 / 1024 / 2048, so
    # the plan filters with the FFT form (fan weight, Parker table and angular
    # weight fused into its input load; npad 512 as a 1024-point transform)
    _say("fft-form fan plan", arc, short, nb)
    cls = sg.SinoFanArc if arc else sg.SinoFanFlat
    kw = dict(nb=nb, d=300.0 / nb, dsd=400.0, dod=150.0)
    rg = cls.short(na=200, **kw) if short else cls(na=200, **kw)
    ig = sg.ImageGeom((96, 96), (1.0, 1.0))
    ig = ig.with_mask(ig.circle(rg.rfov()))
    ell = sg.phantom.random_ellipses(7, 6, 90.0)
    sino = sg.phantom.sino_ellipses(ell, rg).numpy().astype(np.float32)
    plan = sg.plan_fbp(rg, ig, window=sg.Window(sg.Hann()))
    assert len(plan.filter) == 1 << nb.bit_length()
    got = sg.fbp(plan, sino)
    ar = rg.ar(f32=True).astype(np.float32)
    vw = orc.angle_weight(ar) * (orc.parker_fan_short(rg.gamma(), ar, rg.gamma_max())
                                 if short else 1.0)
    sw = orc.fan_sino_weight(rg.gamma(), rg.dso, rg.dsd)
    xc, yc = ig.axes()
    args = (plan.filter, vw, sw, ar, rg.dsd, rg.dso, 0.0, arc, rg.d, 0.0, xc, yc, ig.mask)
    e32 = rel_rmse(got, orc.fbp_fan(sino, *args, dtype=np.float32))
    et = rel_rmse(got, sg.fbp(_toeplitz_plan(sg, rg, ig, window=sg.Window(sg.Hann())), sino))
    _say(f"  rel-RMSE vs Float32 oracle {e32:.3g}, vs the Toeplitz form {et:.3g}")
    assert e32 < TOL32
    # the Toeplitz form's FP32 sums over 1300 bins carry ~5e-6 of their own
    # (the FFT form is the closer of the two to the oracle there)
    assert et < (3e-6 if nb < 1024 else TOL32)


@pytest.mark.parametrize("arc", [False, True])
@pytest.mark.parametrize("short", [False, True])
@pytest.mark.parametrize("ns", [400, 600, 1100])
def test_fft_filter_fdk_plan_vs_oracle(sg, orc, torch_cuda, arc, short, ns):
    # 400 / 600 / 1100-bin cone detector: npad = 512 / 1024 / 2048, the FFT form with the
    # cylinder weight (flat or arc) and the short-scan Parker table fused
    _say("fft-form fdk plan", arc, short, ns)
    cls = sg.CtFanArc if arc else sg.CtFanFlat
    kw = dict(ns=ns, nt=24, ds=300.0 / ns, dt=1.2)
    rg = cls.short(na=59, **kw) if short else cls(na=32, **kw)
    ig = sg.ImageGeom((30, 28, 16), (1.0, 1.0, 1.0))
    ell = np.array([[2.0, 3.0, 1.0, 8.0, 7.0, 6.0, 1.0], [-4.0, 1.0, -1.0, 3.0, 4.0, 5.0, 0.5]])
    proj = sg.phantom.proj_ellipsoids(ell, rg).numpy().astype(np.float32)
    plan = sg.plan_fbp(rg, ig)
    assert len(plan.filter) == 1 << ns.bit_length()
    got = sg.fdk(plan, proj)
    ar = rg.ar(f32=True).astype(np.float32)
    vw = orc.fdk_weight_cyl(rg.s(), rg.t(), rg.dsd, rg.dso, arc)[:, :, None] * \
        orc.angle_weight(ar)
    if short:
        vw = vw * orc.parker_fan_short(rg.gamma(), ar, rg.gamma_max())[:, None, :]
    xc, yc, zc = ig.axes()
    args = (plan.filter, vw, ar, rg.dsd, rg.dso, rg.ds, rg.dt, 0.0, 0.0, arc, xc, yc, zc,
            ig.mask)
    e32 = rel_rmse(got, orc.fdk(proj, *args, dtype=np.float32))
    ref64 = orc.fdk(proj.astype(np.float64), *args)
    et = rel_rmse(got, sg.fdk(_toeplitz_plan(sg, rg, ig), proj))
    e16 = rel_rmse(sg.fdk(sg.plan_fbp(rg, ig, precision="f16s"), proj), ref64)
    _say(f"  rel-RMSE vs Float32 oracle {e32:.3g}, vs the Toeplitz form {et:.3g}, "
         f"F16S vs Float64 {e16:.3g}")
    assert e32 < TOL32
    assert et < (3e-6 if ns < 1024 else TOL32)
    # F16S: the FFT form's scaled FP16 output.  Its error is the FP16 storage
    # rounding averaged over only 32-59 views here (about 1e-4; 3.4e-5 over
    # the 360 views of the 256^2 x 128 workload)
    assert e16 < 2.5e-4
    # BF16: the FFT form's bfloat16 store (round to nearest even) against the
    # oracle's Float32 filtered data, element by element: within half a BF16
    # ulp of the FP32 value plus the FP32 filter's own error
    pb = sg.plan_fbp(rg, ig, precision="bf16")
    gb = sg.fdk(pb, proj)
    stored, scale = pb.read_filtered()
    assert np.all(scale == 1.0)
    qref = orc.sino_filter(proj * vw.astype(np.float32), plan.filter, dtype=np.float32)
    qr = _bf16_round(qref)
    tol = 2.0 ** -8 * np.abs(qref) + 4e-6 * np.abs(qref).max()
    assert np.all(np.abs(stored - qref) <= tol)
    # round to nearest even on the same values: where the FP32 filter's own
    # error cannot move a value across a rounding midpoint (values above 1 %
    # of the largest), the stored values are the oracle's, rounded
    big = np.abs(qref) > 1e-2 * np.abs(qref).max()
    agree = float(np.mean(stored[big] == qr[big]))
    eb = rel_rmse(gb, ref64)
    _say(f"  BF16 store: {agree:.4f} of the values above 1 % of the maximum equal the "
         f"oracle's, rounded; BF16 vs Float64 {eb:.3g}")
    assert agree > 0.99
    assert eb < 2e-3


def test_execute_async_on_current_stream(sg, torch_cuda):
    # Device arrays on torch's current stream return without a host wait;
    # several executes queued back to back must equal the synchronous host
    # path, and a call made while another stream is current still waits.
    import torch
    _say("async executes")
    rg = sg.CtFanFlat(ns=64, nt=40, ds=1.0, dt=1.2, na=32)
    ig = sg.ImageGeom((30, 28, 26), (1.0, 1.0, 1.0))
    rng = np.random.default_rng(21)
    p = np.asfortranarray(rng.standard_normal(rg.dims()).astype(np.float32))
    dev = torch.from_numpy(np.ascontiguousarray(p.transpose(2, 1, 0))).cuda().permute(2, 1, 0)
    plan = sg.plan_fbp(rg, ig)
    host = plan.execute(p)
    outs = [plan.execute(dev) for _ in range(4)]
    for o in outs:
        assert np.array_equal(o.cpu().numpy(), host)
    # An input produced on another stream, behind a delay, with no manual wait:
    # the context's stream must wait for that stream's queued work.
    side = torch.cuda.Stream()
    late = torch.zeros_like(dev)
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        torch.cuda._sleep(20_000_000)          # keeps `side` busy for several ms
        late.copy_(dev)
        o2 = plan.execute(late)
        got = o2.cpu().numpy()
    assert np.array_equal(got, host)


def test_parker_device_vs_oracle(sg, orc, torch_cuda):
    _say("parker")
    rg = sg.SinoFanArc.short(nb=888, d=1.0239, na=984, dsd=949.0, dod=408.0)
    pw = sg.parker_weight(rg)
    ref = orc.parker_fan_short(rg.gamma(), rg.ar(f32=True), rg.gamma_max())
    assert pw.shape == (888, 642)
    assert np.abs(pw - ref).max() < 1e-6
    ct = sg.CtFanFlat.short(ns=64, nt=10, na=100)
    assert sg.parker_weight(ct).shape == (64, 1, ct.na)


# --- fused plans (fbp / fdk) -----------------------------------------------------------------

def _fan_case(sg, arc, short, n=96, nb=120):
    cls = sg.SinoFanArc if arc else sg.SinoFanFlat
    kw = dict(nb=nb, d=1.0, dsd=400.0, dod=150.0)
    rg = cls.short(na=200, **kw) if short else cls(na=200, **kw)
    ig = sg.ImageGeom((n, n), (1.0, 1.0))
    return rg, ig.with_mask(ig.circle(rg.rfov()))


@pytest.mark.parametrize("arc", [True, False])
@pytest.mark.parametrize("short", [False, True])
def test_fbp_fan_plan_vs_oracle(sg, orc, torch_cuda, arc, short):
    _say("fbp fan plan", arc, short)
    rg, ig = _fan_case(sg, arc, short)
    ell = sg.phantom.random_ellipses(5, 6, 96.0)
    sino = sg.phantom.sino_ellipses(ell, rg).numpy().astype(np.float32)
    plan = sg.plan_fbp(rg, ig, window=sg.Window(sg.Hann()))
    got = sg.fbp(plan, sino)
    ar = rg.ar(f32=True).astype(np.float32)
    vw = orc.angle_weight(ar) * (orc.parker_fan_short(rg.gamma(), ar, rg.gamma_max())
                                 if short else 1.0)
    sw = orc.fan_sino_weight(rg.gamma(), rg.dso, rg.dsd)
    xc, yc = ig.axes()
    args = (plan.filter, vw, sw, ar, rg.dsd, rg.dso, 0.0, arc, rg.d, 0.0, xc, yc, ig.mask)
    ref32 = orc.fbp_fan(sino, *args, dtype=np.float32)
    ref64 = orc.fbp_fan(sino.astype(np.float64), *args)
    e32, e64 = rel_rmse(got, ref32), rel_rmse(got, ref64)
    _say(f"  rel-RMSE vs Float32 oracle {e32:.3g}, vs Float64 oracle {e64:.3g}")
    assert e32 < TOL32
    # applied pre-weights equal the reference's sino weight .* view_weight
    w = plan.applied_weights()[:, 0, :]
    assert np.allclose(w, sw[:, None] * vw, rtol=2e-6, atol=1e-9)
    # stack path: recon3[:, :, 1] == recon (test/fbp2/fbp-fan.jl:36-38)
    r3 = sg.fbp(plan, sino.reshape(sino.shape + (1,), order="F"))
    assert np.array_equal(r3[:, :, 0], got)


def test_fbp_par_plan_vs_oracle_and_disk_area(sg, orc, torch_cuda):
    _say("fbp par plan")
    rg = sg.SinoPar(nb=367, d=1.0, na=360)
    ig = sg.ImageGeom((256, 256), (1.0, 1.0))
    ell = sg.phantom.random_ellipses(0, 10, 256.0)
    sino = sg.phantom.sino_ellipses(ell, rg).numpy().astype(np.float32)
    plan = sg.plan_fbp(rg, ig, window=sg.Window(sg.Hann()))
    got = sg.fbp(plan, sino)
    ar = rg.ar(f32=True).astype(np.float32)
    xc, yc = ig.axes()
    args = (plan.filter, orc.angle_weight(ar), ar, rg.d, 0.0, xc, yc, ig.mask)
    e32 = rel_rmse(got, orc.fbp_par(sino, *args, dtype=np.float32))
    e64 = rel_rmse(got, orc.fbp_par(sino.astype(np.float64), *args))
    _say(f"  rel-RMSE vs Float32 oracle {e32:.3g}, vs Float64 oracle {e64:.3g}")
    assert e32 < TOL32
    # test/fbp2/fbp-par.jl:12-33 through the convenience form
    nr, na = 64, 128
    dr = 10 / nr
    r = (np.arange(nr) - (nr - 1) / 2) * dr
    phi = np.arange(na) / na * math.pi
    tau = (r[:, None] - (2 * np.cos(phi) + 1 * np.sin(phi))[None, :]) / 2
    s = (2 * np.where(np.abs(tau) < 1, 2 * np.sqrt(np.clip(1 - tau ** 2, 0, None)), 0)
         ).astype(np.float32)
    recon = sg.fbp(s, dr=dr)
    area = recon.sum() * dr * dr
    assert abs(area - 4 * math.pi) / (4 * math.pi) < 2e-4
    # stack of two slices (5x the second), fbp-par.jl:44-49
    s3 = np.stack([s, 5 * s], axis=2)
    rg2 = sg.SinoPar(nb=nr, na=na, d=dr)
    ig2 = sg.ImageGeom((nr, nr), (dr, dr))
    ig2 = ig2.with_mask(ig2.circle(rg2.rfov()))
    r3 = sg.fbp(sg.plan_fbp(rg2, ig2), s3)
    areas = r3.sum(axis=(0, 1)) * dr * dr / np.array([1, 5])
    assert np.all(np.abs(areas - 4 * math.pi) / (4 * math.pi) < 5e-4)


@pytest.mark.parametrize("arc", [False, True])
@pytest.mark.parametrize("short", [False, True])
def test_fdk_plan_vs_oracle_and_volume(sg, orc, torch_cuda, arc, short):
    # test/fbp3/fdk.jl:14-34 geometry and tolerances
    _say("fdk plan", arc, short)
    cls = sg.CtFanArc if arc else sg.CtFanFlat
    kw = dict(ns=64, nt=40, ds=1.0, dt=1.2)
    rg = cls.short(na=59, **kw) if short else cls(na=32, **kw)
    ig = sg.ImageGeom((30, 28, 26), (1.0, 1.0, 1.0))
    ell = np.array([[2.0, 3.0, 4.0, 8.0, 7.0, 6.0, 1.0]])
    proj = sg.phantom.proj_ellipsoids(ell, rg).numpy().astype(np.float32)
    plan = sg.plan_fbp(rg, ig)
    got = sg.fdk(plan, proj)
    vt = 4 / 3 * math.pi * 8 * 7 * 6
    assert abs(got.sum() - vt) / vt < (3e-3 if not short else 2e-2)
    ar = rg.ar(f32=True).astype(np.float32)
    vw = orc.fdk_weight_cyl(rg.s(), rg.t(), rg.dsd, rg.dso, arc)[:, :, None] * \
        orc.angle_weight(ar)
    if short:
        vw = vw * orc.parker_fan_short(rg.gamma(), ar, rg.gamma_max())[:, None, :]
    xc, yc, zc = ig.axes()
    args = (plan.filter, vw, ar, rg.dsd, rg.dso, rg.ds, rg.dt, 0.0, 0.0, arc, xc, yc, zc,
            ig.mask)
    e32 = rel_rmse(got, orc.fdk(proj, *args, dtype=np.float32))
    e64 = rel_rmse(got, orc.fdk(proj.astype(np.float64), *args))
    _say(f"  rel-RMSE vs Float32 oracle {e32:.3g}, vs Float64 oracle {e64:.3g}")
    assert e32 < TOL32
    w = plan.applied_weights()
    assert np.allclose(w, np.broadcast_to(vw, w.shape), rtol=2e-6, atol=1e-9)


@pytest.mark.parametrize("arc", [False, True])
@pytest.mark.parametrize("short", [False, True])
def test_fdk_volume_from_ray_projections(sg, torch_cuda, arc, short):
    # test/fbp3/fdk.jl:14-34 as written there: projections are line integrals
    # along rays(rg), in the flat case with the reference's polar-angle sign
    _say("fdk volume from rays(rg)", arc, short)
    cls = sg.CtFanArc if arc else sg.CtFanFlat
    kw = dict(ns=64, nt=40, ds=1.0, dt=1.2)
    rg = cls.short(na=59, **kw) if short else cls(na=32, **kw)
    ig = sg.ImageGeom((30, 28, 26), (1.0, 1.0, 1.0))
    ell = np.array([[2.0, 3.0, 4.0, 8.0, 7.0, 6.0, 1.0]])
    proj = sg.raygeom.radon_ellipsoids(*sg.rays(rg), ell).numpy().astype(np.float32)
    got = sg.fdk(sg.plan_fbp(rg, ig), proj)
    vt = 4 / 3 * math.pi * 8 * 7 * 6
    err = abs(got.sum() - vt) / vt
    _say(f"  volume error {err:.3g}")
    assert err < (3e-3 if not short else 2e-2)


def test_fdk_16bit_storage_forms_error(sg, orc, torch_cuda):
    _say("fdk 16-bit storage plans")
    rg = sg.CtFanFlat(ns=128, nt=96, ds=2.4, dt=2.4, na=90, dsd=1200.0, dod=400.0)
    ig = sg.ImageGeom((64, 64, 32), (2.0, 2.0, 2.0))
    ell = sg.phantom.random_ellipsoids(3, 12, 128.0)
    proj = sg.phantom.proj_ellipsoids(ell, rg).numpy().astype(np.float32)
    ar = rg.ar(f32=True).astype(np.float32)
    vw = orc.fdk_weight_cyl(rg.s(), rg.t(), rg.dsd, rg.dso, False)[:, :, None] * \
        orc.angle_weight(ar)
    xc, yc, zc = ig.axes()
    p32 = sg.plan_fbp(rg, ig)
    ref64 = orc.fdk(proj.astype(np.float64), p32.filter, vw, ar, rg.dsd, rg.dso, rg.ds, rg.dt,
                    0.0, 0.0, False, xc, yc, zc, ig.mask)
    errs = {}
    for prec in ("fp32", "bf16", "f16s"):
        plan = p32 if prec == "fp32" else sg.plan_fbp(rg, ig, precision=prec)
        errs[prec] = rel_rmse(sg.fdk(plan, proj), ref64)
    _say("  FDK rel-RMSE vs Float64:", {k: f"{v:.3g}" for k, v in errs.items()})
    assert errs["fp32"] < 1e-5
    assert errs["f16s"] < 1e-4          # the 16-bit form that meets the 1e-4 bar
    assert errs["bf16"] < 5e-3          # BF16 storage: recorded, above 1e-4


@pytest.mark.parametrize("prec", ["fp32", "f16s", "bf16"])
def test_fdk_stored_data_exact_vs_oracle(sg, orc, torch_cuda, prec):
    # The backprojection reads exactly the stored filtered data: widen the
    # plan's stored values exactly, apply the per-view scale and backproject
    # them with the Float32 oracle.
    # The z axis is symmetric, so the staged kernel's mirror form would serve
    # the lower half from mirrored windows; the FP16 form's pair model below
    # (upper tap a + fp16(b - a)) is the non-mirror kernel's, so that form is
    # checked with the mirror form off, and the mirror form after it.
    _say("fdk stored-data parity", prec)
    rg = sg.CtFanArc(ns=66, nt=48, ds=1.0, dt=1.2, na=40, dsd=300.0, dod=100.0)
    ig = sg.ImageGeom((24, 24, 16))
    rng = np.random.default_rng(7)
    proj = rng.standard_normal(rg.dims()).astype(np.float32)
    ctx = sg._lib.context(0)
    ctx.cbct_mirror = "never"
    try:
        plan = sg.plan_fbp(rg, ig, precision=prec)
    finally:
        ctx.cbct_mirror = "fp16"
    got = sg.fdk(plan, proj)
    q, scale = plan.read_filtered()
    raw = plan.read_stored()
    assert raw.shape == (68, 48, 40)                  # rows padded to a multiple of 4
    assert np.all(raw[66:] == 0)
    if prec == "f16s":
        assert np.all(np.frexp(scale)[0] == 0.5)      # powers of two
        assert np.abs(q).max() <= 2.0 ** 14
    xc, yc, zc = ig.axes()
    if prec == "f16s":
        # the staged kernel keeps (value, row difference) pairs in FP16: the
        # upper tap is a + fp16(b - a); check against the pair-reading oracle
        a = q.astype(np.float16)
        d = np.zeros_like(a)
        d[:, :-1, :] = a[:, 1:, :] - a[:, :-1, :]          # FP16 arithmetic, rounded once
        sc = scale[None, None, :].astype(np.float64)
        ref = orc.cbct_back_pairs(a.astype(np.float64) * sc, d.astype(np.float64) * sc,
                                  rg.ar(f32=True), rg.dsd, rg.dso, rg.ds, rg.dt, True,
                                  xc, yc, zc, ig.mask, dtype=np.float32)
        # The mirror form reads the lower half's upper taps as stored values
        # (b, not a + fp16(b - a)): against the exact interpolation of the
        # stored values it is at least as close as the pair model.
        exact = orc.cbct_back(a.astype(np.float64) * sc, rg.ar(f32=True), rg.dsd, rg.dso,
                              rg.ds, rg.dt, 0.0, 0.0, True, xc, yc, zc, ig.mask)
        pm = sg.plan_fbp(rg, ig, precision=prec)
        assert sg.tuning(pm)["cbct_kernel"] == "staged_mirror"       # the form's default
        got_m = sg.fdk(pm, proj)
        e_m, e_p = rel_rmse(got_m, exact), rel_rmse(got, exact)
        _say(f"  vs the exact interpolation of the stored values: mirror {e_m:.3g}, "
             f"pair model {e_p:.3g}")
        assert e_m <= 1.05 * e_p + 1e-6
    else:
        q = q.astype(np.float64) * scale[None, None, :]
        ref = orc.cbct_back(q, rg.ar(f32=True), rg.dsd, rg.dso, rg.ds, rg.dt, 0.0, 0.0, True,
                            xc, yc, zc, ig.mask, dtype=np.float32)
    assert rel_rmse(got, ref) < TOL32


@pytest.mark.parametrize("case", ["flat", "arc_opts_edges"])
def test_fdk_f16s_packed16_vs_fp32_interp(sg, orc, torch_cuda, case):
    # The packed-FP16 interpolation of the scaled-FP16 form (cone-beam kernel
    # "staged_packed16") against the default FP32 interpolation of the same
    # stored data and against Float64.  Its extra rounding (position to 1/1024
    # row, FP16 interpolated values and weights) about doubles the FP16
    # storage error (6.7e-5 against 3.4e-5 on the 256^2 x 128 FDK workload).
    # "arc_opts_edges": a helical arc detector with the exact t magnification
    # and a detector short enough that the volume's top and bottom leave it
    # (windows that are not interior, lanes on the gather path).
    _say("fdk f16s packed16", case)
    if case == "flat":
        rg = sg.CtFanFlat(ns=128, nt=96, ds=2.4, dt=2.4, na=90, dsd=1200.0, dod=400.0)
        ig = sg.ImageGeom((64, 64, 32), (2.0, 2.0, 2.0))
        ell = sg.phantom.random_ellipsoids(3, 12, 128.0)
        opts = {}
    else:
        rg = sg.CtFanArc(ns=96, nt=30, ds=1.5, dt=1.5, na=72,
                         src=sg.CtSourceHelix(pitch=0.3, source_z0=-2.0))
        ig = sg.ImageGeom((48, 44, 40), (1.5, 1.5, 1.5))
        ell = sg.phantom.random_ellipsoids(4, 10, 64.0)
        opts = dict(use_source_zs=True, arc_t_exact=True)
    proj = sg.phantom.proj_ellipsoids(ell, rg).numpy().astype(np.float32)
    ctx = sg._lib.context(0)
    old = ctx.cbct_kernel
    try:
        ctx.cbct_kernel = "staged_packed16"
        packed = sg.fdk(sg.plan_fbp(rg, ig, precision="f16s", **opts), proj)
    finally:
        ctx.cbct_kernel = old
    f16 = sg.fdk(sg.plan_fbp(rg, ig, precision="f16s", **opts), proj)
    f32 = sg.fdk(sg.plan_fbp(rg, ig, **opts), proj)
    d = rel_rmse(packed, f16)
    e_pk, e16 = rel_rmse(packed, f32), rel_rmse(f16, f32)
    _say(f"  packed vs FP32-interpolated {d:.3g}; from the FP32 plan: packed {e_pk:.3g}, "
         f"FP32-interpolated {e16:.3g}")
    assert d < 1e-3
    assert e_pk < 2.5 * e16 + 1e-6
    if case == "flat":
        ar = rg.ar(f32=True).astype(np.float32)
        vw = orc.fdk_weight_cyl(rg.s(), rg.t(), rg.dsd, rg.dso, False)[:, :, None] * \
            orc.angle_weight(ar)
        xc, yc, zc = ig.axes()
        ref64 = orc.fdk(proj.astype(np.float64), sg.plan_fbp(rg, ig).filter, vw, ar, rg.dsd,
                        rg.dso, rg.ds, rg.dt, 0.0, 0.0, False, xc, yc, zc, ig.mask)
        e64 = rel_rmse(packed, ref64)
        _say(f"  packed vs Float64 {e64:.3g}")
        assert e64 < 2.5e-4


# --- K6 distance-driven projector ------------------------------------------------------

@pytest.mark.parametrize("nPix,nDet,na,f32", [(32, 57, 24, False), (128, 226, 90, True),
                                              (64, 101, 360, False)])
def test_project_bdd_vs_oracle(sg, orc, torch_cuda, nPix, nDet, na, f32):
    _say("project_bdd", nPix, nDet, na, f32)
    geo = dict(DSD=949.0, DS0=541.0, pSize=512.0 / nPix, dSize=1.0239 * 910 / nDet,
               nPix=nPix, nDet=nDet, angle=np.arange(na) * 2 * math.pi / na)
    rng = np.random.default_rng(40)
    x = rng.random((nPix, nPix)) - 0.3
    if f32:
        x = x.astype(np.float32)
    got = sg.project_bdd(x, geo)
    ref = orc.project_bdd(x.astype(np.float64), **geo)
    assert got.shape == (na, nDet) and got.dtype == np.float64
    assert rel_rmse(got, ref) < 1e-12
    # device arrays in, device array out
    xt = torch_cuda.from_numpy(np.ascontiguousarray(x.T)).cuda().T
    got_d = sg.project_bdd(xt, geo).cpu().numpy()
    assert np.allclose(got_d, got, rtol=0, atol=1e-12 * np.abs(got).max())


@pytest.mark.parametrize("nPix,nDet,na,a0", [(32, 57, 24, 0.0), (128, 226, 90, 0.0),
                                             (40, 57, 24, math.pi / 2), (50, 80, 18, 3.5)])
def test_backproject_bdd_vs_oracle_and_adjoint(sg, orc, torch_cuda, nPix, nDet, na, a0):
    # a0 = pi / 2: the first view maps on the y axis, so the reference's
    # first-view aliasing adds it in the other frame; 40 and 50 pixels leave
    # partial 16 x 16 tiles at the image edge.
    _say("backproject_bdd", nPix, nDet, na, a0)
    geo = dict(DSD=949.0, DS0=541.0, pSize=512.0 / nPix, dSize=1.0239 * 910 / nDet,
               nPix=nPix, nDet=nDet, angle=(a0 + np.arange(na) * 2 * math.pi / na) % (2 * math.pi))
    rng = np.random.default_rng(41)
    y = rng.random((na, nDet))
    for alias in (True, False):
        got = sg.backproject_bdd(y, geo, ref_alias=alias)
        ref = orc.backproject_bdd(y, **geo, ref_alias=alias)
        assert rel_rmse(got, ref) < 1e-12
        # owner-computes: no atomics, the same bits every call
        assert np.array_equal(sg.backproject_bdd(y, geo, ref_alias=alias), got)
    # adjoint identity of the GPU pair (without the reference's aliasing quirk)
    x = rng.random((nPix, nPix))
    lhs = np.dot(sg.project_bdd(x, geo).ravel(), y.ravel())
    rhs = np.dot(x.ravel(), sg.backproject_bdd(y, geo, ref_alias=False).ravel())
    assert lhs == pytest.approx(rhs, rel=1e-12)


# --- K7 batches (one HIP graph) and autotuning ----------------------------------------

def _colmajor(torch_cuda, a):
    a = np.asarray(a, dtype=np.float32)
    return torch_cuda.from_numpy(np.ascontiguousarray(a.T)).cuda().permute(
        *range(a.ndim - 1, -1, -1))


def _batch_specs(sg):
    """Uneven reconstructions: 2-D parallel / fan (arc, flat) slices and FDK
    volumes (flat, arc; FP32 and scaled FP16)."""
    specs = []
    for kind, n, na in [("arc", 64, 90), ("flat", 96, 120), ("par", 80, 60), ("arc", 128, 180),
                        ("arc", 64, 90)]:
        if kind == "par":
            rg = sg.SinoPar(nb=int(1.45 * n), na=na, d=1.0)
        else:
            rg = (sg.SinoFanArc if kind == "arc" else sg.SinoFanFlat)(
                nb=int(1.6 * n), na=na, d=1.0, dsd=4.0 * n, dod=1.5 * n)
        specs.append((rg, sg.ImageGeom((n, n), (1.0, 1.0)), "fp32"))
    for kind, n, na, prec in [("flat", 32, 60, "fp32"), ("arc", 48, 72, "f16s"),
                              ("flat", 40, 48, "fp32")]:
        rg = (sg.CtFanArc if kind == "arc" else sg.CtFanFlat)(
            ns=2 * n, nt=int(1.5 * n), ds=1.0, dt=1.0, na=na, dsd=6.0 * n, dod=2.0 * n)
        specs.append((rg, sg.ImageGeom((n, n, n // 2), (1.0, 1.0, 1.0)), prec))
    return specs


def test_batch_graph_matches_unbatched(sg, torch_cuda):
    _say("batch graph vs unbatched")
    rng = np.random.default_rng(21)
    specs = _batch_specs(sg)
    plans = [sg.plan_fbp(rg, ig, precision=prec) for rg, ig, prec in specs]
    # the last plan gets a second item (items of one plan share its buffers)
    items = [(p, spec) for p, spec in zip(plans, specs)] + [(plans[-1], specs[-1])]
    sinos = [_colmajor(torch_cuda, rng.standard_normal(spec[0].dims())) for _, spec in items]
    outs = [sg.recon._empty_colmajor(spec[1].dims, torch_cuda.device("cuda:0"))
            for _, spec in items]
    batch = sg.Batch()
    for (p, _), x, y in zip(items, sinos, outs):
        batch.add(p, x, y)
    batch.execute()                       # autotune + capture (both forms timed) + replay
    info = batch.info()
    assert (info["items"], info["branches"]) == (len(items), 4)
    assert info["timed_ms_queues"] > 0 and info["timed_ms_launches"] > 0
    assert (info["queues"] > 0) == (info["timed_ms_queues"] < info["timed_ms_launches"])
    # in the queue form every item but the two of the shared plan and the
    # gather-tuned cone-beam ones backprojects through a queue launch
    batch.execute(queues=True)
    info = batch.info()
    gathered = sum(1 for p, spec in zip(plans, specs) if isinstance(spec[0], sg.CtFan)
                   and sg.tuning(p)["cbct_kernel"] == "gather" and p is not plans[-1])
    assert info["queued_items"] == len(items) - 2 - gathered and info["queues"] >= 2
    assert all(sg.tuning(p)["tuned"] for p in plans)
    refs = []
    for (p, spec), x in zip(items, sinos):
        r = sg.recon._empty_colmajor(spec[1].dims, torch_cuda.device("cuda:0"))
        p.execute(x, out=r)
        refs.append(r.clone())
    # eager re-executes and graph replays, with other kernels in between, must
    # reproduce the first eager result (a result that depends on what ran
    # before, e.g. LDS read before it is written, shows up here)
    for rep in range(3):
        for i, ((p, spec), x, r) in enumerate(zip(items, sinos, refs)):
            r2 = sg.recon._empty_colmajor(spec[1].dims, torch_cuda.device("cuda:0"))
            p.execute(x, out=r2)
            e = float((r2 - r).norm() / r.norm())
            assert e <= 1e-6, ("eager", rep, i, e)
    for rep in range(6):
        if rep:
            for y in outs:                # eager work between replays
                y.zero_()
            plans[0].execute(sinos[0])
            torch_cuda.cuda.synchronize()
            batch.execute()
        for i, (y, r) in enumerate(zip(outs, refs)):
            e = float((y - r).norm() / r.norm())
            assert e <= 1e-6, ("graph", rep, i, e)
    assert batch.last_ms() > 0
    # the graph of whole plan executes (no queues), then the queues again
    for queues in (False, True):
        for y in outs:
            y.zero_()
        batch.execute(queues=queues)
        assert (batch.info()["queues"] > 0) == queues
        for i, (y, r) in enumerate(zip(outs, refs)):
            e = float((y - r).norm() / r.norm())
            assert e <= 1e-6, ("queues" if queues else "plain", i, e)


def test_autotune_choice_and_cache(sg, orc, torch_cuda):
    _say("autotune")
    rng = np.random.default_rng(22)
    rg = sg.CtFanFlat(ns=96, nt=72, ds=1.0, dt=1.0, na=40, dsd=300.0, dod=100.0)
    ig = sg.ImageGeom((40, 40, 24))
    proj = rng.standard_normal(rg.dims()).astype(np.float32)
    x = _colmajor(torch_cuda, proj)
    dev = torch_cuda.device("cuda:0")
    p1 = sg.plan_fbp(rg, ig)
    y1 = sg.recon._empty_colmajor(ig.dims, dev)
    t = sg.autotune(p1, x, y1)
    assert t["tuned"] and t["cbct_kernel"] in ("staged", "gather", "staged_mirror", "staged_z8",
                                               "staged_z8w")
    ref = orc.fdk(proj.astype(np.float64), p1.filter,
                  orc.fdk_weight_cyl(rg.s(), rg.t(), rg.dsd, rg.dso, False)[:, :, None] *
                  orc.angle_weight(rg.ar(f32=True).astype(np.float32)),
                  rg.ar(f32=True), rg.dsd, rg.dso, rg.ds, rg.dt, 0.0, 0.0, False,
                  *ig.axes(), ig.mask)
    assert rel_rmse(y1.cpu().numpy(), ref) < 1e-5       # the tuned result
    # a plan of the same size bucket takes the cached choice
    p2 = sg.plan_fbp(rg, ig)
    y2 = sg.recon._empty_colmajor(ig.dims, dev)
    assert sg.autotune(p2, x, y2) == t
    assert float((y2 - y1).abs().max()) == 0.0


# --- K8 learned sparse-view correction ------------------------------------------------

def test_net_forward_backward_vs_oracle(sg, torch_cuda):
    from oracle import nn as onn
    from sinograms_amd.learned import ConvCorrection
    _say("net forward/backward vs oracle")
    rng = np.random.default_rng(41)
    layers, ch, R, N = 4, 32, 5, 150
    net = ConvCorrection(layers, ch, seed=3)
    p = (rng.standard_normal(net.num_params) * 0.05).astype(np.float32)
    net.set_params(p)
    assert np.array_equal(net.params(), p)
    q = rng.standard_normal((N, R)).astype(np.float32)          # (n, rows), column-major
    gout = rng.standard_normal((N, R)).astype(np.float32)
    qd = _colmajor(torch_cuda, q)
    out = net.forward(qd, keep=True)
    net.zero_grad()
    net.backward(_colmajor(torch_cuda, gout))
    g = net.grads()
    params = onn.split_params(p.astype(np.float64), layers, ch)
    ref, cache = onn.forward(params, q.T.astype(np.float64))
    gref = onn.flatten(onn.backward(params, cache, gout.T.astype(np.float64)))
    assert rel_rmse(out.cpu().numpy().T, ref) < 1e-6
    assert rel_rmse(g, gref) < 1e-5
    # every layer's weight and bias gradients, not only the total
    o = 0
    for w, b in params:
        for a in (w, b):
            k = a.size
            assert rel_rmse(g[o:o + k], gref[o:o + k]) < 1e-4
            o += k
    # a fresh network is the identity (plain FBP is the starting point)
    net0 = ConvCorrection(layers, ch, seed=4)
    assert float((net0.forward(qd) - qd).abs().max()) == 0.0


@pytest.mark.parametrize("layers", [2, 3, 4, 5])
def test_net_fused_inference(sg, torch_cuda, layers):
    # Inference (keep=False) of a 32-channel network is one fused kernel
    # (activations in LDS); training forwards (keep=True) run layer by layer.
    # Both against the float64 oracle, on rows shorter than one fused tile,
    # spanning several tiles with a ragged end, and after parameter updates
    # (the fused form's BF16 weights are rebuilt when the parameters change).
    from oracle import nn as onn
    from sinograms_amd.learned import ConvCorrection
    _say("fused net inference, layers", layers)
    rng = np.random.default_rng(50 + layers)
    ch = 32
    net = ConvCorrection(layers, ch, seed=5)
    p = (rng.standard_normal(net.num_params) * 0.05).astype(np.float32)
    net.set_params(p)
    for N, R in ((37, 3), (300, 4), (888, 2)):
        q = rng.standard_normal((N, R)).astype(np.float32)
        qd = _colmajor(torch_cuda, q)
        fused = net.forward(qd).cpu().numpy()
        layered = net.forward(qd, keep=True).cpu().numpy()
        ref, _ = onn.forward(onn.split_params(p.astype(np.float64), layers, ch),
                             q.T.astype(np.float64))
        assert rel_rmse(fused.T, ref) < 1e-6, N
        assert rel_rmse(layered.T, ref) < 1e-6, N
        assert rel_rmse(fused, layered) < 1e-6, N
    # a backward pass and an Adam step change the parameters: the next fused
    # forward uses the new ones
    net.zero_grad()
    net.backward(_colmajor(torch_cuda, rng.standard_normal((888, 2)).astype(np.float32)))
    net.adam(1e-2)
    p2 = net.params()
    assert not np.array_equal(p2, p)
    q = rng.standard_normal((300, 4)).astype(np.float32)
    qd = _colmajor(torch_cuda, q)
    ref2, _ = onn.forward(onn.split_params(p2.astype(np.float64), layers, ch),
                          q.T.astype(np.float64))
    assert rel_rmse(net.forward(qd).cpu().numpy().T, ref2) < 1e-6


def test_back_t_is_the_exact_transpose(sg, orc, torch_cuda):
    from sinograms_amd.learned import _empty_colmajor
    _say("back_fan transpose")
    rng = np.random.default_rng(42)
    for rg in (sg.SinoFanFlat(nb=47, na=11, d=1.3, dsd=120.0, dod=40.0),
               sg.SinoFanArc(nb=47, na=11, d=1.3, dsd=120.0, dod=40.0),
               sg.SinoPar(nb=47, na=11, d=1.3)):
        ig = sg.ImageGeom((20, 18), (1.7, 1.7))
        plan = sg.plan_fbp(rg, ig)
        ctx = plan._ctx
        L = sg._lib.load()
        dev = torch_cuda.device("cuda:0")
        # explicit matrix of the backprojection, column by column (unit sinograms)
        nb, na = rg.dims()
        A = np.zeros((ig.dims[0] * ig.dims[1], nb * na))
        e = _empty_colmajor((nb, na), dev)
        img = _empty_colmajor(ig.dims, dev)
        for c in range(nb * na):
            e.zero_()
            e.t().reshape(-1)[c] = 1.0
            ctx.check(L.sgx_plan_back(plan.h, ctypes.c_void_p(e.data_ptr()),
                                      ctypes.c_void_p(img.data_ptr()), 0))
            A[:, c] = img.t().reshape(-1).cpu().numpy()
        r = rng.standard_normal(ig.dims).astype(np.float32)
        qt = _empty_colmajor((nb, na), dev)
        ctx.check(L.sgx_plan_back_t(plan.h, ctypes.c_void_p(_colmajor(torch_cuda, r).data_ptr()),
                                    ctypes.c_void_p(qt.data_ptr()), 0))
        want = A.T @ r.ravel(order="F")
        assert rel_rmse(qt.cpu().numpy().ravel(order="F"), want) < 1e-6


def test_learned_correction_beats_plain_fbp(sg, torch_cuda):
    # A small sparse-view problem: train on distance-driven projections of
    # random-ellipse phantoms, evaluate on held-out phantoms against the
    # dense-view FBP target; the trained correction must beat plain FBP.
    from sinograms_amd import learned
    _say("learned correction training")
    prob = learned.SparseViewProblem(npix=64, pix=4.0, nb=111, d=4.2, na_sparse=24,
                                     na_dense=180)
    net = learned.ConvCorrection(layers=4, channels=16, seed=0)
    before = learned.evaluate(prob, net, range(1000, 1006))
    assert abs(before["psnr_learned"] - before["psnr_plain"]) < 1e-9     # f = 0 at start
    hist = learned.train(prob, net, range(64), batch=8, lr=2e-3, epochs=6)
    after = learned.evaluate(prob, net, range(1000, 1006))
    _say(f"  loss {hist['losses'][0]:.4g} -> {hist['losses'][-1]:.4g}; PSNR plain "
         f"{after['psnr_plain']:.2f} dB, learned {after['psnr_learned']:.2f} dB; SSIM "
         f"{after['ssim_plain']:.4f} -> {after['ssim_learned']:.4f}")
    assert hist["losses"][-1] < hist["losses"][0]
    assert after["psnr_learned"] > after["psnr_plain"] + 0.2


# --- analytic phantom projector ---------------------------------------------------

@pytest.mark.parametrize("kind", ["par", "arc", "flat"])
def test_radon_ellipses_hip_vs_torch(sg, torch_cuda, kind):
    # libsgx's FP64 projector against the torch implementation, whose rays
    # are pinned to the reference's conventions on the CPU
    _say("radon ellipses", kind)
    kw = dict(nb=130, d=1.3, offset=0.25, na=77, orbit_start=11.0)
    rg = {"par": lambda: sg.SinoPar(**kw),
          "arc": lambda: sg.SinoFanArc(dsd=400.0, dod=150.0, **kw),
          "flat": lambda: sg.SinoFanFlat(dsd=400.0, dod=150.0, **kw)}[kind]()
    ell = sg.phantom.random_ellipses(5, 20, 150.0)
    got = sg.phantom.sino_ellipses(ell, rg, device="cuda")
    assert got.dtype == torch_cuda.float64 and tuple(got.shape) == (rg.nb, rg.na)
    ref = sg.phantom.sino_ellipses(ell, rg, device="cpu").numpy()
    assert np.abs(got.cpu().numpy() - ref).max() <= 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("arc", [False, True])
def test_radon_ellipsoids_hip_vs_torch(sg, torch_cuda, arc):
    _say("radon ellipsoids", arc)
    rg = (sg.CtFanArc if arc else sg.CtFanFlat)(
        ns=50, nt=30, ds=1.4, dt=1.1, offset_s=0.3, offset_t=-0.2, na=25, orbit_start=5.0,
        dsd=300.0, dod=100.0, src=sg.CtSourceHelix(pitch=0.7, source_z0=-5.0))
    ell = sg.phantom.random_ellipsoids(6, 10, 60.0)
    for zs in (None, rg.source_zs()):
        for views in (None, [0, 3, 7, 24]):
            got = sg.phantom.proj_ellipsoids(ell, rg, "cuda", views, source_zs=zs)
            ref = sg.phantom.proj_ellipsoids(ell, rg, "cpu", views, source_zs=zs).numpy()
            assert tuple(got.shape) == ref.shape
            assert np.abs(got.cpu().numpy() - ref).max() <= 1e-10 * np.abs(ref).max()


def test_helical_fdk_places_the_object(sg, torch_cuda):
    # helical projections (source and detector raised per view) reconstructed
    # with the per-view source z put an ellipsoid at its height; the
    # reference's geometry (source z ignored) does not
    _say("helical fdk")
    rg = sg.CtFanArc(ns=96, nt=48, ds=1.0, dt=1.0, na=90, dsd=300.0, dod=100.0,
                     src=sg.CtSourceHelix(pitch=0.6, source_z0=-4.0))
    ig = sg.ImageGeom((40, 40, 30), (1.0, 1.0, 1.0))
    ell = np.array([[3.0, -2.0, 2.5, 10.0, 8.0, 6.0, 1.0]])
    proj = sg.phantom.proj_ellipsoids(ell, rg, "cuda", source_zs=rg.source_zs()).float()
    Z = torch_cuda.as_tensor(ig.axes()[2], device="cuda")[None, None, :]
    vt = 4 / 3 * math.pi * 10 * 8 * 6
    res = {}
    for name, kw in (("helix", dict(use_source_zs=True)), ("reference", {})):
        img = sg.fdk(sg.plan_fbp(rg, ig, **kw), proj).double()
        res[name] = (float((img * Z).sum() / img.sum()), abs(float(img.sum()) - vt) / vt)
    _say(f"  z centroid / volume error: {res}")
    assert abs(res["helix"][0] - 2.5) < 0.05 and res["helix"][1] < 0.02
    assert abs(res["reference"][0] - 2.5) > 1.0


def test_batch_recaptures_after_plan_change(sg, torch_cuda):
    # Changing a plan's cone-beam options after a batch captured it frees and
    # replaces the plan's per-view source z table: the next batch execute must
    # capture again (and so follow the new geometry), not replay stale pointers.
    import ctypes
    _say("batch recapture after a plan change")
    rng = np.random.default_rng(23)
    rg = sg.CtFanFlat(ns=64, nt=48, ds=1.0, dt=1.0, na=40, dsd=200.0, dod=70.0,
                      src=sg.CtSourceHelix(pitch=0.5, source_z0=-4.0))
    ig = sg.ImageGeom((32, 32, 24))
    dev = torch_cuda.device("cuda:0")
    x = _colmajor(torch_cuda, rng.standard_normal(rg.dims()))
    plan = sg.plan_fbp(rg, ig)
    y = sg.recon._empty_colmajor(ig.dims, dev)
    batch = sg.Batch()
    batch.add(plan, x, y)
    batch.execute(tune=False)
    first = y.clone()
    zs = np.ascontiguousarray(rg.source_zs(), dtype=np.float64)
    opts = sg._lib.SgxCbctOpts(ctypes.c_void_p(zs.ctypes.data), 0)
    plan._ctx.check(sg._lib.load().sgx_plan_set_cbct_opts(plan.h, ctypes.byref(opts)))
    batch.execute(tune=False)
    ref = sg.fdk(sg.plan_fbp(rg, ig, use_source_zs=True), x)
    assert rel_rmse(y.cpu().numpy(), ref.cpu().numpy()) <= 1e-6
    assert rel_rmse(first.cpu().numpy(), ref.cpu().numpy()) > 1e-3
    plan._ctx.check(sg._lib.load().sgx_plan_set_cbct_opts(plan.h, None))
    batch.execute(tune=False)
    assert rel_rmse(y.cpu().numpy(), first.cpu().numpy()) <= 1e-6


# --- the benchmark workloads, as configured -------------------------------------------

@pytest.mark.parametrize("short", [False, True])
def test_c1_fan_arc_workload(sg, orc, torch_cuda, short):
    # C1 as configured: 888 x 984 (or the 642-view Parker short scan) -> 512^2
    # at 0.98 mm in the rfov circle, noisy.  The Float32 restatement is itself
    # about 1e-5 from exact arithmetic here (tests/test_fp32_floor.py), so the
    # bars are: the GPU's arithmetic error (against Float64 arithmetic on the
    # Float32 path's rounded inputs, FP32 ramp filter included) well below that; at least as close to
    # Float64 as the Float32 restatement; and within the 1e-5 bar of the
    # Float32 restatement, or within its arithmetic error where that is larger.
    from conftest import c1_case, fan_references
    _say("C1 workload", "short" if short else "full")
    rg, ig, sino = c1_case(sg, short)
    plan = sg.plan_fbp(rg, ig)
    dev = _colmajor(torch_cuda, sino)
    got = sg.fbp(plan, dev).cpu().numpy()
    assert np.array_equal(got, sg.fbp(plan, sino))          # host and device pointers
    assert np.all(got[~ig.mask] == 0)
    ref32, ref64, ref64i = fan_references(orc, rg, ig, sino, plan.filter)
    arith = rel_rmse(got, ref64i)
    e32 = rel_rmse(got, ref32)
    e64 = rel_rmse(got, ref64)
    floor = rel_rmse(ref32, ref64i)
    _say(f"  GPU: arithmetic {arith:.3g}, vs Float32 oracle {e32:.3g}, vs Float64 {e64:.3g}; "
         f"Float32 oracle arithmetic {floor:.3g}, vs Float64 {rel_rmse(ref32, ref64):.3g}")
    assert arith < 5e-6
    assert e64 <= rel_rmse(ref32, ref64)
    assert e32 <= max(TOL32, floor + arith)


def test_c2_fdk_workload(sg, orc, torch_cuda):
    # C2 as configured: CtFanFlat 512 x 384 at 1.2 mm, dso 800, dsd 1200, 360
    # views -> 256 x 256 x 128 at 1 mm, 12 random ellipsoids (seed 0), in the
    # three storage forms.  Bars: FP32 within 1e-5 of the Float32 oracle; the
    # scaled-FP16 form within the north-star 1e-4 of Float64; the BF16 form's
    # storage rounding (an 8-bit significand) leaves it above 1e-4, checked
    # here against its measured ceiling.
    _say("C2 workload")
    rg = sg.CtFanFlat(ns=512, nt=384, ds=1.2, dt=1.2, na=360, dsd=1200.0, dod=400.0)
    ig = sg.ImageGeom((256, 256, 128), (1.0, 1.0, 1.0))
    ell = sg.phantom.random_ellipsoids(0, 12, 256.0)
    proj = sg.phantom.proj_ellipsoids(ell, rg, device=torch_cuda.device("cuda:0")).float()
    vols = {}
    for prec in ("fp32", "f16s", "bf16"):
        plan = sg.plan_fbp(rg, ig, precision=prec)
        vols[prec] = sg.fdk(plan, proj).cpu().numpy()
        Hk = plan.filter
        del plan
    ph = proj.double().cpu().numpy()                      # (ns, nt, na), column-major view
    ar = rg.ar(f32=True).astype(np.float32)
    vw = orc.fdk_weight_cyl(rg.s(), rg.t(), rg.dsd, rg.dso, False)[:, :, None] * \
        orc.angle_weight(ar)
    xc, yc, zc = ig.axes()
    args = (Hk, vw, ar, rg.dsd, rg.dso, rg.ds, rg.dt, 0.0, 0.0, False, xc, yc, zc, ig.mask)
    ref32 = orc.fdk(ph.astype(np.float32), *args, dtype=np.float32)
    ref64 = orc.fdk(ph, *args)
    e = {f: (rel_rmse(v, ref32), rel_rmse(v, ref64)) for f, v in vols.items()}
    _say("  " + ", ".join(f"{f}: vs Float32 {a:.3g}, vs Float64 {b:.3g}" for f, (a, b) in e.items()))
    assert e["fp32"][0] < TOL32
    assert e["f16s"][1] <= 1e-4
    assert e["bf16"][1] < 3e-4


def test_c3_batch_workload(sg, torch_cuda):
    # C3 as configured: the 48-reconstruction mix of seed 3 (32 arc fan slices
    # of 128-512 px and 180-720 views, 16 FDK volumes of 64-192 voxels and
    # 120-360 views) as one autotuned graph, each result within rel-RMSE 1e-6
    # of the unbatched GPU path.
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    _say("C3 workload")
    dev = torch_cuda.device("cuda:0")
    items = bench.c3_items(sg, dev, seed=3)
    assert len(items) == 48
    plans = [sg.plan_fbp(rg, ig) for rg, ig, _ in items]
    outs = [sg.recon._empty_colmajor(ig.dims, dev) for _, ig, _ in items]
    batch = sg.Batch()
    for p, (_, _, x), y in zip(plans, items, outs):
        batch.add(p, x, y)
    batch.execute()
    batch.execute()                       # a replay
    worst = 0.0
    for p, (rg, ig, x), y in zip(plans, items, outs):
        r = p.execute(x)
        worst = max(worst, float((y - r).norm() / r.norm()))
    _say(f"  worst rel-RMSE vs unbatched {worst:.3g}")
    assert worst <= 1e-6


def test_c4_learned_workload(sg, torch_cuda):
    # C4 as configured: the committed correction network (4 layers, 32
    # channels, width 9) on the 512^2 sparse-view problem (90 short-scan views
    # of the 888-bin fan): its forward equals the float64 oracle network on a
    # filtered sinogram, and on held-out phantoms it beats plain sparse-view
    # FBP (the fallback) against the dense-view FBP target.
    from oracle import nn as onn
    from sinograms_amd import learned
    _say("C4 workload")
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "sinograms.jl_amd", "data", "sparse_view_correction_fan90.npy")
    params = np.load(path)
    prob = learned.SparseViewProblem()
    net = learned.ConvCorrection(layers=4, channels=32)
    net.set_params(params)
    sino, _ = prob.sample(10 ** 6)
    q = prob.filter(sino)
    out = net.forward(q).cpu().numpy()
    qn = q.cpu().numpy()
    ref, _ = onn.forward(onn.split_params(params.astype(np.float64), 4, 32),
                         qn.T.astype(np.float64))
    e = rel_rmse(out.T, ref)
    r = learned.evaluate(prob, net, range(10 ** 6, 10 ** 6 + 8))
    _say(f"  forward vs oracle {e:.3g}; PSNR plain {r['psnr_plain']:.2f} dB, learned "
         f"{r['psnr_learned']:.2f} dB; SSIM {r['ssim_plain']:.3f} -> {r['ssim_learned']:.3f}")
    assert e < 1e-6
    assert r["psnr_learned"] > r["psnr_plain"] + 3.0
    assert r["ssim_learned"] > r["ssim_plain"] + 0.1


@pytest.mark.parametrize("arc", [False, True])
@pytest.mark.parametrize("exact", [False, True])
def test_cbct_back_source_offset_vs_oracle(sg, orc, torch_cuda, cbct_kernel, arc, exact):
    # the lateral source offset the reference keeps commented out
    # (cbct-back.jl:178), pinned on the CPU to its fan-beam formula; with the
    # exact arc t magnification and a helix in the second case
    _say("cbct_back source offset", cbct_kernel, arc, exact)
    rng = np.random.default_rng(14)
    src = sg.CtSourceHelix(pitch=0.8, source_z0=-6.0) if exact else sg.CtSourceCircle()
    rg = (sg.CtFanArc if arc else sg.CtFanFlat)(ns=48, nt=36, ds=1.1, dt=1.3, na=30,
                                                 offset_s=0.25, offset_t=-0.5, dsd=120.0,
                                                 dod=50.0, source_offset=3.3, src=src)
    proj = rng.standard_normal(rg.dims()).astype(np.float32)
    ig = sg.ImageGeom((29, 23, 26), (1.0, 0.9, 1.1), (0.3, 0.0, -0.2))
    xc, yc, zc = ig.axes()
    for ia_skip in (1, 2):
        got = sg.cbct_back(proj, rg, ig, ia_skip=ia_skip, use_source_offset=True,
                           use_source_zs=exact, arc_t_exact=exact)
        ref = orc.cbct_back(proj, rg.ar(f32=True), rg.dsd, rg.dso, rg.ds, rg.dt, rg.offset_s,
                            rg.offset_t, arc, xc, yc, zc, ig.mask, ia_skip=ia_skip,
                            dtype=np.float32, source_zs=rg.source_zs() if exact else None,
                            arc_t_exact=exact, source_offset=rg.source_offset)
        assert rel_rmse(got, ref) < TOL32, ia_skip
    base = sg.cbct_back(proj, rg, ig, ia_skip=2, use_source_zs=exact, arc_t_exact=exact)
    assert rel_rmse(base, got) > 1e-3


@pytest.mark.parametrize("arc", [False, True])
def test_fdk_plan_source_offset_vs_oracle(sg, orc, torch_cuda, arc):
    _say("fdk plan source offset", arc)
    rng = np.random.default_rng(15)
    rg = (sg.CtFanArc if arc else sg.CtFanFlat)(ns=64, nt=40, ds=1.0, dt=1.2, na=32,
                                                 source_offset=-2.5)
    ig = sg.ImageGeom((30, 28, 26), (1.0, 1.0, 1.0))
    proj = rng.standard_normal(rg.dims()).astype(np.float32)
    plan = sg.plan_fbp(rg, ig, use_source_offset=True)
    got = sg.fdk(plan, proj)
    ar = rg.ar(f32=True).astype(np.float32)
    vw = orc.fdk_weight_cyl(rg.s(), rg.t(), rg.dsd, rg.dso, arc)[:, :, None] * \
        orc.angle_weight(ar)
    xc, yc, zc = ig.axes()
    ref = orc.fdk(proj, plan.filter, vw, ar, rg.dsd, rg.dso, rg.ds, rg.dt, 0.0, 0.0, arc, xc,
                  yc, zc, ig.mask, dtype=np.float32, source_offset=rg.source_offset)
    assert rel_rmse(got, ref) < TOL32
    assert rel_rmse(sg.fdk(sg.plan_fbp(rg, ig), proj), got) > 1e-3


@pytest.mark.parametrize("window", [None, "hann"])
def test_mojette_plan_vs_oracle(sg, orc, torch_cuda, window):
    # plan_fbp(rg::SinoMoj; how=:mojette) + fbp against the oracle's
    # restatement, which builds each view's ramp for its own spacing
    _say("mojette plan", window)
    d, nb, na = 1.0, 128, 200
    rg = sg.SinoMoj(nb=nb, d=d, na=na, offset=0.25)
    ig = sg.ImageGeom((96, 90), (d, d))
    ig = ig.with_mask(ig.circle(rg.rfov()))
    ell = sg.phantom.random_ellipses(9, 6, 60.0)
    sino = sg.phantom.sino_ellipses(ell, rg).numpy().astype(np.float32)
    win = sg.Window(sg.Hann()) if window else None
    plan = sg.plan_fbp(rg, ig, how="mojette", window=win)
    assert isinstance(plan, sg.MojPlan)
    got = sg.fbp(plan, sino)
    xc, yc = ig.axes()
    w = orc.fbp_window("hann", 256) if window else None
    ref = orc.fbp_moj(sino.astype(np.float64), rg.ar(f32=True), d, rg.offset, xc, yc, ig.mask,
                      npad=256, window=w)
    e = rel_rmse(got, ref)
    _say(f"  rel-RMSE vs the Float64 oracle {e:.3g}")
    assert e < TOL32
    assert np.all(got[~ig.mask] == 0)
    # the default plan of a mojette geometry is the same
    assert np.array_equal(sg.fbp(sg.plan_fbp(rg, ig, window=win), sino), got)
