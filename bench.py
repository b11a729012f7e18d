"""Benchmark: FDK (cone-beam, flat panel) on one MI355X.

Workload (the configuration the headline metric is quoted on):
  CtFanFlat, 512 x 384 detector at 1.2 mm, dso = 800 mm, dsd = 1200 mm,
  360 views over 360 deg; 256 x 256 x 128 volume at 1 mm; 12 random
  ellipsoids (seed 0) with analytic cone-beam projections (Float64 -> Float32).

A step is one fdk(plan, proj): fused weights + ramp filter (FFT form) +
voxel-driven backprojection, with the projections already resident in HBM.
value is voxel-view updates per second of the whole step for the plain FP32
form; the 16-bit storage forms are reported beside it.  The roofline object
describes the dominant kernel (the backprojector).  Its taps come from LDS,
so it is bound by instruction issue: `frac` is its VALU issue rate (rocprofv3
instruction counts of the committed summary over the launch time measured
here) against the MI355X issue ceiling; its HBM traffic, measured and
compulsory, is reported beside it, and the gather-equivalent model figure
(16 B per voxel-view update for FP32 storage) only as a labelled model.

The 2-D configurations (parallel beam; arc fan beam, full and short scan)
are reported beside it under "slices", the distance-driven projector pair
under "dd_projector".

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batch] [--no-host]
                       [--no-slices]
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "sinograms.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = ("FDK backproj Gupdates/s & % HBM peak; mixed vs plain FP32 at "
          "rel-RMSE≤1e-4")
HBM_PEAK_GBS = 8000.0
MU_UNIT = 0.005           # 1 / mm: attenuation of a unit phantom density (C1 noise)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_problem(device):
    import numpy as np
    import sinograms_amd as sg
    rg = sg.CtFanFlat(ns=512, nt=384, ds=1.2, dt=1.2, na=360, dsd=1200.0, dod=400.0)
    ig = sg.ImageGeom((256, 256, 128), (1.0, 1.0, 1.0))
    fov = 256.0
    ell = sg.phantom.random_ellipsoids(0, 12, fov)
    proj = sg.phantom.proj_ellipsoids(ell, rg, device=device).float()   # (ns, nt, na) col-major
    return sg, rg, ig, ell, proj


# rocprofv3 counter summaries (text, under profiles/).  Each holds lines
#   counters <kernel> <source-tag> <sha16> key=value ...
# with per-launch means (valu_insts, lds_insts, fetch_bytes, write_bytes)
# and the hash of the kernel sources the counters were taken from.
COUNTER_FILES = ("profiles/fdk256_counters.md", "profiles/slices_counters.md")
KERNEL_SOURCES = {
    "k_cbct": ("sinograms.jl_amd/csrc/k_cbct.hip", "sinograms.jl_amd/csrc/sgx_internal.h"),
    "k_back2d": ("sinograms.jl_amd/csrc/k_back2d.hip", "sinograms.jl_amd/csrc/sgx_internal.h"),
    "k_filter": ("sinograms.jl_amd/csrc/k_filter.hip", "sinograms.jl_amd/csrc/sgx_internal.h"),
    "k_bdd": ("sinograms.jl_amd/csrc/k_bdd.hip", "sinograms.jl_amd/csrc/sgx_internal.h"),
    "k_nn": ("sinograms.jl_amd/csrc/k_nn.hip", "sinograms.jl_amd/csrc/sgx_internal.h"),
}
# VALU issue ceiling of an MI355X: 256 CUs x 4 SIMDs, one wave64 VALU
# instruction per SIMD every 4 cycles (16 lanes per cycle), 2.4 GHz.
VALU_PEAK_G = 256 * 4 * 2.4e9 / 4 / 1e9


def source_hash(tag):
    """sha256 (first 16 hex digits) of the sources a kernel is built from."""
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES[tag]:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def counters(kernel):
    """The committed per-launch counters of the first kernel whose name
    starts with `kernel` ({} when no summary lists one), with "current" True
    when the summary was taken from the kernel sources in this tree.  A name
    may carry a workload tag after "@" (two workloads launching one grid of
    one kernel, told apart in the summary); `kernel` "k@tag" matches the
    tagged line only."""
    kernel, _, tag = kernel.partition("@")
    for name in COUNTER_FILES:
        try:
            with open(os.path.join(ROOT, name)) as f:
                for line in f:
                    parts = line.split()
                    if len(parts) >= 4 and parts[0] == "counters" and \
                            parts[1].startswith(kernel) and \
                            (not tag or parts[1].endswith("@" + tag)):
                        r = {k: float(v) for k, v in (kv.split("=") for kv in parts[4:])}
                        r["source"] = name
                        r["current"] = parts[3] == source_hash(parts[2])
                        return r
        except OSError:
            continue
    return {}


def roofline(kernel, kernel_ms, compulsory_bytes, units, unit_name, model_bytes_per_unit=None):
    """Roofline object of one launch of `kernel` (average time kernel_ms,
    measured in this run with HIP events).

    The staged backprojectors serve their taps from LDS, so they are bound by
    instruction issue, not by HBM: the binding resource is VALU issue, and
    `achieved` = VALU wave instructions per launch (rocprofv3 SQ_INSTS_VALU,
    from the committed summary) over the measured launch time, against the
    issue ceiling.  HBM is reported beside it, measured (FETCH_SIZE /
    WRITE_SIZE of the summary, calibrated) and compulsory (every input and
    output byte once).  When the summary was taken from other kernel sources
    than these, its counts are not used: the object falls back to the
    compulsory HBM bytes over the measured time."""
    ks = kernel_ms * 1e-3
    c = counters(kernel)
    hbm = {"compulsory_bytes": compulsory_bytes,
           "compulsory_GBs": compulsory_bytes / ks / 1e9,
           "frac_compulsory": compulsory_bytes / ks / 1e9 / HBM_PEAK_GBS, "peak_GBs": HBM_PEAK_GBS}
    r = {"kernel": kernel, "kernel_ms": kernel_ms}
    if c.get("current") and "valu_insts" in c:
        ach = c["valu_insts"] / ks / 1e9
        traffic = c.get("fetch_bytes", 0.0) + c.get("write_bytes", 0.0) or None
        r.update({"bound": "valu", "achieved": ach, "peak": VALU_PEAK_G,
                  "unit": "G VALU wave-instructions/s", "frac": ach / VALU_PEAK_G,
                  "traffic": traffic, "counters_source": c["source"],
                  "valu_insts_per_launch": c["valu_insts"],
                  f"valu_per_{unit_name}": c["valu_insts"] * 64.0 / units})
        if "lds_insts" in c:
            r["lds_insts_per_launch"] = c["lds_insts"]
        if traffic:
            hbm.update({"measured_bytes": traffic, "measured_GBs": traffic / ks / 1e9,
                        "frac_measured": traffic / ks / 1e9 / HBM_PEAK_GBS})
    else:
        r.update({"bound": "hbm", "achieved": hbm["compulsory_GBs"], "peak": HBM_PEAK_GBS,
                  "unit": "GB/s", "frac": hbm["frac_compulsory"], "traffic": None,
                  "note": "no counter summary of these kernel sources: compulsory bytes only"})
    r["hbm"] = hbm
    if model_bytes_per_unit:
        # the gather-equivalent figure (taps x stored bytes per update): a
        # model of the data the update reads, most of it from LDS -- not HBM
        r["gather_equivalent_GBs"] = model_bytes_per_unit * units / ks / 1e9
    return r


def time_steps(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def filter_form(plan):
    """Which ramp-filter kernel a plan runs: the FFT form at npad = 512, 1024
    or 2048, else the Toeplitz form on MFMA (also when SGX_FILTER_TOEPLITZ=1 was
    set while the plan was made)."""
    return "toeplitz" if len(plan.filter) not in (512, 1024, 2048) or \
        getattr(plan, "_toeplitz", False) else "fft"


def filter_figures(plan, fm, nb, ncols, out_bytes):
    """Filter figures: HBM bytes moved per launch (input read once, output
    written once) over the launch time, and for the Toeplitz form the
    algorithmic 2 nb^2 ncols flops over it."""
    byts = 4.0 * nb * ncols + out_bytes * nb * ncols
    r = {"filter_form": filter_form(plan), "filter_hbm_GBs": byts / (fm * 1e-3) / 1e9}
    if r["filter_form"] == "toeplitz":
        r["filter_TFLOPs_algorithmic"] = 2.0 * nb * nb * ncols / (fm * 1e-3) / 1e12
    return r


def plan_toeplitz(sg, rg, ig, **kw):
    """A plan whose filter is the Toeplitz form (the A/B against the FFT form)."""
    os.environ["SGX_FILTER_TOEPLITZ"] = "1"
    try:
        plan = sg.plan_fbp(rg, ig, **kw)
    finally:
        del os.environ["SGX_FILTER_TOEPLITZ"]
    plan._toeplitz = True
    return plan


def filter_forms_leg(sg, device, reps):
    """The plan's filter kernel alone (HIP events around it), the FFT form
    against the Toeplitz form, for flat-panel cone-beam plans with a 512-bin
    (npad 1024) and a 1500-bin (npad 2048) detector, 128 rows x 360 views
    (46080 columns), onto a small volume."""
    import torch
    out = {}
    for ns in (512, 1500):
        rg = sg.CtFanFlat(ns=ns, nt=128, ds=600.0 / ns, dt=1.2, na=360, dsd=1200.0,
                          dod=400.0)
        ig = sg.ImageGeom((32, 32, 8), (4.0, 4.0, 4.0))
        proj = torch.rand(rg.na, rg.nt, rg.ns, device=device).permute(2, 1, 0)
        y = sg.recon._empty_colmajor(ig.dims, device)
        r = {"npad": None}
        for form in ("fft", "toeplitz"):
            plan = plan_toeplitz(sg, rg, ig) if form == "toeplitz" else sg.plan_fbp(rg, ig)
            r["npad"] = len(plan.filter)
            plan.execute(proj, out=y)
            fm, _ = kernel_times(plan, proj, y, reps)
            r[form + "_ms"] = fm
            assert filter_form(plan) == form
            del plan
        r["speedup"] = r["toeplitz_ms"] / r["fft_ms"]
        r["hbm_GBs_fft"] = 8.0 * ns * rg.nt * rg.na / (r["fft_ms"] * 1e-3) / 1e9
        out[f"CtFanFlat_{ns}x128x360"] = r
        log(f"  filter ns={ns} npad={r['npad']}: FFT {r['fft_ms']:.3f} ms, "
            f"Toeplitz {r['toeplitz_ms']:.3f} ms")
    return out


def event_ms(fn, n):
    """Average milliseconds of n calls of fn, from HIP events recorded on
    torch's current stream (the stream the calls run on)."""
    import torch
    fn()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(n):
        fn()
    b.record()
    b.synchronize()
    return a.elapsed_time(b) / n


def kernel_times(plan, proj, out, n):
    """Average (filter_ms, back_ms) over n executes, from HIP events recorded
    on the stream the kernels run on."""
    f = b = 0.0
    for _ in range(n):
        plan.execute(proj, out=out)
        fm, bm = plan.times()
        f += fm
        b += bm
    return f / n, b / n


def c3_items(sg, device, seed=3):
    """The batching workload: 48 independent reconstructions of uneven size
    drawn from a seed -- 32 fan-beam arc slices (image side in {128, 256,
    384, 512}, views in {180, 360, 720}; the 888-bin arc detector at dso 541
    / dsd 949) and 16 FDK volumes (side in {64, 128, 192}, 120-360 views; the
    512 x 384 flat panel at dso 800 / dsd 1200), random-ellipse / ellipsoid
    phantoms with analytic projections on the device.  Returns [(rg, ig, x)]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    items = []
    for i in range(32):
        n = int(rng.choice([128, 256, 384, 512]))
        na = int(rng.choice([180, 360, 720]))
        rg = sg.SinoFanArc(nb=888, na=na, d=1.0239, dsd=949.0, dod=949.0 - 541.0)
        ig = sg.ImageGeom((n, n), (0.98 * 512 / n, 0.98 * 512 / n))
        ell = sg.phantom.random_ellipses(seed * 1000 + i, 15, 0.98 * 512)
        x = sg.phantom.sino_ellipses(ell, rg, device=device).float().T.contiguous().T
        items.append((rg, ig, x))
    for i in range(16):
        n = int(rng.choice([64, 128, 192]))
        na = int(rng.integers(120, 361))
        rg = sg.CtFanFlat(ns=512, nt=384, ds=1.2, dt=1.2, na=na, dsd=1200.0, dod=400.0)
        ig = sg.ImageGeom((n, n, n), (256.0 / n,) * 3)
        ell = sg.phantom.random_ellipsoids(seed * 1000 + 100 + i, 12, 256.0)
        items.append((rg, ig, sg.phantom.proj_ellipsoids(ell, rg, device=device).float()))
    return items


def batch_leg(sg, device, steps, seed=3):
    """The batching form: the 48 reconstructions of c3_items run as one
    autotuned HIP graph, timed against the same executes issued one after
    another.  Two graph forms: queue launches (filters on parallel branches,
    then one launch per backprojector variant over a size-ordered table of
    tile / voxel-block tasks) and whole plan executes on parallel branches;
    the capture times both and keeps the faster, and both are reported."""
    import numpy as np
    import torch
    items = c3_items(sg, device, seed)
    updates = sum(int(np.prod(ig.dims)) * rg.na for rg, ig, _ in items)
    plans = [sg.plan_fbp(rg, ig) for rg, ig, _ in items]
    outs = [sg.recon._empty_colmajor(ig.dims, device) for _, ig, _ in items]
    batch = sg.Batch()
    for p, (_, _, x), y in zip(plans, items, outs):
        batch.add(p, x, y)
    batch.execute()                                  # autotune + capture + first replay
    refs = [sg.recon._empty_colmajor(ig.dims, device) for _, ig, _ in items]

    def serial():
        for p, (_, _, x), r in zip(plans, items, refs):
            p.execute(x, out=r)
    t_serial = time_steps(serial, steps, 1)
    forced = {}
    for q in (False, True):
        batch.execute(queues=q)
        forced[q] = time_steps(lambda: batch.execute(queues=q), steps, 1)
        forced[q, "err"] = max(float((y - r).norm() / r.norm()) for y, r in zip(outs, refs))
        if q:
            qinfo = batch.info()
    batch.execute()                                  # capture with the timed choice
    info = batch.info()
    t_graph = time_steps(batch.execute, steps, 1)
    err = max(float((y - r).norm() / r.norm()) for y, r in zip(outs, refs))
    tunes = [sg.tuning(p) for p in plans]
    return {"workload": "48 reconstructions of uneven size (seed 3): 32 fan-beam arc slices "
                        "128-512 px x 180-720 views, 16 FDK volumes 64-192 voxels x 120-360 views",
            "items": len(items), "updates": updates, "branches": info["branches"],
            "chosen_form": "queues" if info["queues"] else "launches",
            "capture_timed_ms": {"queues": info["timed_ms_queues"],
                                 "launches": info["timed_ms_launches"]},
            "graph_ms": t_graph * 1e3, "serial_ms": t_serial * 1e3,
            "speedup_vs_serial": t_serial / t_graph,
            "queue_form": {"ms": forced[True] * 1e3, "speedup_vs_serial": t_serial / forced[True],
                           "queues": qinfo["queues"], "queued_items": qinfo["queued_items"],
                           "tasks": qinfo["tasks"],
                           "max_rel_rmse_vs_unbatched": forced[True, "err"]},
            "launch_form": {"ms": forced[False] * 1e3,
                            "speedup_vs_serial": t_serial / forced[False],
                            "max_rel_rmse_vs_unbatched": forced[False, "err"]},
            "value": updates / t_graph / 1e9, "unit": "Gupdates/s",
            "max_rel_rmse_vs_unbatched": err,
            "cbct_kernels": sorted({t["cbct_kernel"] for t, (rg, _, _) in zip(tunes, items)
                                    if isinstance(rg, sg.CtFan)}),
            "b2d_shapes": sorted({(t["want_blocks"], t["view_chunk"])
                                  for t, (rg, _, _) in zip(tunes, items)
                                  if not isinstance(rg, sg.CtFan)})}


def cpu_leg(sg, rg, ig, proj, Hk, vols, nslab):
    """Oracle timing on a bounded sample and the accuracy of every GPU form on
    it.  Sample: the nslab central z-slices of the volume (all 256 x 256
    columns, all 360 views)."""
    import numpy as np
    import oracle
    oracle.lib()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ar = rg.ar(f32=True).astype(np.float32)
    xc, yc, zc = ig.axes()
    z0 = ig.dims[2] // 2 - nslab // 2
    zs = zc[z0:z0 + nslab]
    mask = np.ones((len(xc), len(yc), nslab), bool)
    p = proj.double().cpu().permute(2, 1, 0).numpy().transpose(2, 1, 0)   # (ns, nt, na) F-order
    vw = oracle.fdk_weight_cyl(rg.s(), rg.t(), rg.dsd, rg.dso, False)[:, :, None] * \
        oracle.angle_weight(ar)
    # Float64 oracle: weight, filter (direct Toeplitz), backproject
    t = time.perf_counter()
    q64 = oracle.sino_filter(p * vw, Hk.astype(np.float64))
    t_filter64 = time.perf_counter() - t
    ref64 = oracle.cbct_back(q64, ar, rg.dsd, rg.dso, rg.ds, rg.dt, 0.0, 0.0, False, xc, yc, zs,
                             mask)
    # reference-precision (Float32) oracle; its backprojection is the timed baseline
    q32 = oracle.sino_filter((p * vw).astype(np.float32), Hk, dtype=np.float32)
    t = time.perf_counter()
    ref32 = oracle.cbct_back(q32, ar, rg.dsd, rg.dso, rg.ds, rg.dt, 0.0, 0.0, False, xc, yc,
                             zs, mask, dtype=np.float32)
    t_back32 = time.perf_counter() - t
    updates = ref32.size * rg.na

    def err(x, r):
        return float(np.sqrt(((x - r) ** 2).sum() / (r ** 2).sum()))
    g = {f: v[:, :, z0:z0 + nslab] for f, v in vols.items()}
    acc = {
        "sample": f"z-slices {z0}..{z0 + nslab - 1} of 128 (all 256x256 columns, 360 views)",
        "fp32_vs_f32_oracle": err(g["fp32"], ref32),
        "fp32_vs_f64_oracle": err(g["fp32"], ref64),
        "f16s_vs_f64_oracle": err(g["f16s"], ref64),
        "bf16_vs_f64_oracle": err(g["bf16"], ref64),
        "f16s_packed16_vs_f64_oracle": err(g["f16s_pk"], ref64) if "f16s_pk" in g else None,
        "f32_oracle_vs_f64_oracle": err(ref32, ref64),
    }
    cpu = {
        "value": updates / t_back32 / 1e9,
        "unit": "Gupdates/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle Float32 cbct_back (C++/OpenMP restatement, not the reference) "
                   f"on {nslab} of 128 z-slices: {updates} voxel-view updates in "
                   f"{t_back32:.2f} s; CPU {cpu_model()}; filter excluded "
                   f"(Float64 direct filter took {t_filter64:.1f} s)"),
    }
    return cpu, acc


def slice_legs(sg, device, steps, warmup, with_cpu):
    """The 2-D configurations, each one plan_fbp + fbp with the sinogram
    resident in HBM: C0, parallel beam (367 bins at 1 mm, 360 views over 180
    deg, Hann window) onto 256 x 256 at 1 mm, 10 random ellipses (seed 0); C1,
    the 888-bin arc fan (1.0239 mm, dso 541, dsd 949) onto 512 x 512 at
    0.98 mm inside the circle of radius rfov, 15 random ellipses (seed 1) with
    1e5-photon Poisson noise (seed 2; densities taken in units of 0.005 / mm,
    so the longest ray attenuates by about e^-4.3; counts of 0 clamped to 1),
    as a full scan (984 views over 360 deg) and a short scan (642 views over
    180 deg + fan angle, Parker weights).  Analytic Float64 sinograms cast to Float32.
    The backprojector's roofline is its VALU issue rate (see roofline()),
    with the gather-equivalent 8 B per update (2 taps x 4 B) beside it as a model.  With `with_cpu`, the oracle's Float32 backprojection is
    timed on the host and every result is compared with the Float64 and
    Float32 oracles."""
    import numpy as np
    import torch
    cases = []
    rg0 = sg.SinoPar(nb=367, d=1.0, na=360)
    ig0 = sg.ImageGeom((256, 256), (1.0, 1.0))
    s0 = sg.phantom.sino_ellipses(sg.phantom.random_ellipses(0, 10, 256.0), rg0).numpy()
    cases.append(("C0_par", "SinoPar 367 bins x 360 views over 180 deg, Hann window -> "
                  "256x256 at 1 mm, full mask; 10 random ellipses (seed 0)",
                  rg0, ig0, s0.astype(np.float32), sg.Window(sg.Hann())))
    kw = dict(nb=888, d=1.0239, dsd=949.0, dod=949.0 - 541.0)
    full = sg.SinoFanArc(na=984, **kw)
    short = sg.SinoFanArc(na=642, orbit=full.orbit_short(), **kw)
    ig1 = sg.ImageGeom((512, 512), (0.98, 0.98))
    ig1 = ig1.with_mask(ig1.circle(full.rfov()))
    ell1 = sg.phantom.random_ellipses(1, 15, 512 * 0.98)
    for name, rg in (("C1_fan_arc_full", full), ("C1_fan_arc_short", short)):
        p = sg.phantom.sino_ellipses(ell1, rg).numpy()
        # densities read in units of MU_UNIT for the photon statistics (the
        # longest ray then attenuates by about e^-4.3); the noisy line
        # integrals return to phantom units
        noisy = (sg.phantom.poisson_log(MU_UNIT * p, 1e5, 2) / MU_UNIT).astype(np.float32)
        cases.append((name, f"SinoFanArc 888 bins x {rg.na} views over {rg.orbit:.2f} deg "
                      "(dso 541, dsd 949) -> 512x512 at 0.98 mm in the rfov circle "
                      f"({int(ig1.mask.sum())} pixels); 15 random ellipses (seed 1), "
                      "1e5-photon Poisson noise (seed 2, densities in units of 0.005/mm, "
                      "zero counts clamped to 1)",
                      rg, ig1, noisy, None))
    out = {}
    for name, desc, rg, ig, sino, win in cases:
        log(f"timing {name} ...")
        plan = sg.plan_fbp(rg, ig, window=win) if win is not None else sg.plan_fbp(rg, ig)
        x = torch.from_numpy(np.ascontiguousarray(sino.T)).to(device).T      # (nb, na) col-major
        y = sg.recon._empty_colmajor(ig.dims, device)
        t = time_steps(lambda: plan.execute(x, out=y), steps, warmup)
        fm, bm = kernel_times(plan, x, y, max(3, min(steps, 10)))
        updates = int(ig.mask.sum()) * rg.na
        r = {"workload": desc, "updates_per_step": updates, "ms_per_step": t * 1e3,
             "value": updates / t / 1e9, "unit": "Gupdates/s", "projections_per_s": rg.na / t,
             "back_kernel_ms": bm, "filter_ms": fm,
             "roofline": roofline(back2d_kernel_name(rg) + SCAN_TAG.get(name, ""), bm,
                                  4 * rg.nb * rg.na + 4 * int(np.prod(ig.dims)), updates,
                                  "update", 8.0)}
        r.update(filter_figures(plan, fm, rg.nb, rg.na, 4))
        log(f"  {name}: step {t * 1e3:.3f} ms (filter {fm:.3f} ms, back {bm:.3f} ms)")
        if with_cpu:
            r.update(_slice_cpu(sg, rg, ig, sino, plan, y.cpu().numpy()))
        out[name] = r
    return out


# the C1 scans launch one grid of one kernel; the counter summary tags them
SCAN_TAG = {"C1_fan_arc_full": "@full", "C1_fan_arc_short": "@short"}


def back2d_kernel_name(rg):
    """The 2-D backprojector a plan of `rg` launches (csrc/k_back2d.hip), as
    a kernel-name prefix: parallel beam, or the arc fan's tile-relative form
    (its series length is chosen per launch)."""
    import sinograms_amd as sg
    return "back2d_kernel<0," if isinstance(rg, sg.SinoPar) else "back2d_kernel<4,"


def phantom_leg(sg, device, steps):
    """The analytic projector (FP64, one thread per ray): the FDK workload's
    projections (12 ellipsoids onto 512 x 384 x 360 flat-detector rays) and
    the 888 x 984 arc fan sinogram of 15 ellipses, by libsgx's kernels and, for
    comparison, by the torch elementwise form on the same GPU."""
    import numpy as np
    from sinograms_amd import phantom as ph
    rg3 = sg.CtFanFlat(ns=512, nt=384, ds=1.2, dt=1.2, na=360, dsd=1200.0, dod=400.0)
    ell3 = ph.random_ellipsoids(0, 12, 256.0)
    rg2 = sg.SinoFanArc(nb=888, na=984, d=1.0239, dsd=949.0, dod=949.0 - 541.0)
    ell2 = ph.random_ellipses(1, 15, 500.0)
    idx = np.arange(rg3.na)
    cases = {
        "cone_512x384x360_12_ellipsoids": (
            rg3.ns * rg3.nt * rg3.na, len(ell3),
            lambda: ph.proj_ellipsoids(ell3, rg3, device),
            lambda: ph._proj_ellipsoids_torch(ell3, rg3, device, idx, None, 64)),
        "fan_888x984_15_ellipses": (
            rg2.nb * rg2.na, len(ell2),
            lambda: ph.sino_ellipses(ell2, rg2, device),
            lambda: ph._sino_ellipses_torch(ell2, rg2, device)),
    }
    out = {}
    for name, (rays, nc, hip, tch) in cases.items():
        th = time_steps(hip, steps, 1)
        tt = time_steps(tch, max(1, steps // 3), 1)
        out[name] = {"hip_ms": th * 1e3, "torch_ms": tt * 1e3, "rays_per_s": rays / th,
                     "ray_components_per_s": rays * nc / th, "dtype": "f64"}
        log(f"  analytic projector {name}: {th * 1e3:.3f} ms (torch form {tt * 1e3:.3f} ms)")
    return out


def dd_leg(sg, device, steps):
    """The distance-driven projector pair (FP64, as the reference computes
    it) at the learned form's size: a 512 x 512 image of 0.98 mm onto 984
    views x 888 bins of 1.0239 mm (dso 541, dsd 949), and at the reference
    test's size (128 x 128 onto 90 x 226).  Device arrays in and out."""
    import numpy as np
    import torch
    out = {}
    for nPix, nDet, na in [(512, 888, 984), (128, 226, 90)]:
        geo = dict(DSD=949.0, DS0=541.0, pSize=512.0 * 0.98 / nPix,
                   dSize=1.0239 * 888 / nDet, nPix=nPix, nDet=nDet,
                   angle=np.arange(na) * 2 * np.pi / na)
        g = torch.Generator(device=device).manual_seed(0)
        img = torch.rand((nPix, nPix), device=device, dtype=torch.float64,
                         generator=g).t().contiguous().t()
        sino = torch.rand((na, nDet), device=device, dtype=torch.float64,
                          generator=g).t().contiguous().t()
        tp = time_steps(lambda: sg.project_bdd(img, geo), steps, 2)
        tb = time_steps(lambda: sg.backproject_bdd(sino, geo), steps, 2)
        r = {"project_ms": tp * 1e3, "backproject_ms": tb * 1e3,
             "rays_per_s": na * nDet / tp, "pixel_views_per_s": nPix * nPix * na / tp,
             "dtype": "f64"}
        if nPix == 512:
            # both calls run on torch's current stream: HIP events around
            # each (the main kernel and a short finishing kernel)
            units = nPix * nPix * na
            io = 8 * (nPix * nPix + na * nDet)
            cp = event_ms(lambda: sg.project_bdd(img, geo), steps)
            cb = event_ms(lambda: sg.backproject_bdd(sino, geo), steps)
            r["roofline"] = roofline("bdd_project_kernel<double>", cp, io, units, "update")
            r["roofline"]["note_time"] = "HIP events around the call (bdd_project_kernel + bdd_finish_kernel)"
            r["roofline_adjoint"] = roofline("bdd_back_kernel", cb, io, units, "update")
            r["roofline_adjoint"]["note_time"] = ("HIP events around the call "
                                                  "(bdd_back_table_kernel + bdd_back_kernel)")
        out[f"{nPix}x{nPix}_to_{na}x{nDet}"] = r
        log(f"  project_bdd {nPix}^2 -> {na} x {nDet}: {tp * 1e3:.3f} ms, "
            f"adjoint {tb * 1e3:.3f} ms")
    return out


def learned_leg(sg, device, steps):
    """The learned sparse-view correction (C4): the committed network (4
    layers, 32 channels, width 9) on the 512 x 512 problem of 90 short-scan
    views of the 888-bin flat fan.  Quality on 8 held-out phantoms (PSNR and
    SSIM of plain and corrected sparse-view FBP against the dense-view FBP),
    inference time (the fused kernel on one 888 x 90 filtered sinogram, and
    the corrected reconstruction against the plain one) and training
    throughput (batch 8, 16 phantoms sampled beforehand, forward, backward
    and Adam on the device)."""
    import numpy as np
    import torch
    from sinograms_amd import learned
    path = os.path.join(ROOT, "sinograms.jl_amd", "data", "sparse_view_correction_fan90.npy")
    params = np.load(path)
    prob = learned.SparseViewProblem()
    net = learned.ConvCorrection(layers=4, channels=32)
    net.set_params(params)
    log("  quality on 8 held-out phantoms ...")
    quality = learned.evaluate(prob, net, range(10 ** 6, 10 ** 6 + 8))
    sino, _ = prob.sample(10 ** 6)
    q = prob.filter(sino)
    out = net.forward(q)
    t_inf = time_steps(lambda: net.forward(q, out=out), steps, 3)
    t_plain = time_steps(lambda: prob.reconstruct(sino), steps, 3)
    t_learned = time_steps(lambda: prob.reconstruct(sino, net), steps, 3)
    nb, na = prob.rg_sparse.dims()
    flop = 2.0 * nb * na * (9 * 32 + 2 * 9 * 32 * 32 + 9 * 32)
    log("  training (batch 8) ...")
    tnet = learned.ConvCorrection(layers=4, channels=32)
    tnet.set_params(params)
    hist = learned.train(prob, tnet, range(16), batch=8, lr=1e-4, epochs=2)
    r = {"workload": "4-layer 32-channel width-9 correction on 90 short-scan views x 888 bins "
                     "(flat fan, dso 541, dsd 949) -> 512 x 512 at 0.98 mm; committed weights",
         "psnr_plain_db": quality["psnr_plain"], "psnr_learned_db": quality["psnr_learned"],
         "ssim_plain": quality["ssim_plain"], "ssim_learned": quality["ssim_learned"],
         "eval_phantoms": 8,
         "inference_ms": t_inf * 1e3, "inference_tflops": flop / t_inf / 1e12,
         "recon_plain_ms": t_plain * 1e3, "recon_learned_ms": t_learned * 1e3,
         "correction_overhead_ms": (t_learned - t_plain) * 1e3,
         "train_steps": hist["steps"], "train_batch": 8,
         "train_ms_per_step": hist["train_s"] / max(hist["steps"], 1) * 1e3,
         "train_phantoms_per_s": hist["phantoms_per_s"]}
    log(f"  learned: PSNR {r['psnr_plain_db']:.2f} -> {r['psnr_learned_db']:.2f} dB, inference "
        f"{r['inference_ms']:.3f} ms, overhead {r['correction_overhead_ms']:.3f} ms, training "
        f"{r['train_ms_per_step']:.2f} ms per batch-8 step")
    return r


def _slice_cpu(sg, rg, ig, sino, plan, got):
    """Oracle timing (Float32 backprojection of the filtered sinogram) and the
    GPU result's rel-RMSE against the Float32 and Float64 oracles."""
    import numpy as np
    import oracle
    oracle.lib()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ar = rg.ar(f32=True).astype(np.float32)
    xc, yc = ig.axes()
    vw = oracle.angle_weight(ar)
    if isinstance(rg, sg.SinoPar):
        args = (plan.filter, vw, ar, rg.d, 0.0, xc, yc, ig.mask)
        run = oracle.fbp_par
    else:
        if rg.orbit < 360:
            vw = vw * oracle.parker_fan_short(rg.gamma(), ar, rg.gamma_max())
        sw = oracle.fan_sino_weight(rg.gamma(), rg.dso, rg.dsd)
        args = (plan.filter, vw, sw, ar, rg.dsd, rg.dso, 0.0, True, rg.d, 0.0, xc, yc, ig.mask)
        run = oracle.fbp_fan
    ref64 = run(sino.astype(np.float64), *args)
    ref32 = run(sino, *args, dtype=np.float32)
    # timed baseline: the Float32 backprojection alone, on its own filtered input
    s = sino.astype(np.float32) * np.asarray(vw, np.float32)
    if not isinstance(rg, sg.SinoPar):
        s = s * np.asarray(sw, np.float32)[:, None]
    q = oracle.sino_filter(s, plan.filter, dtype=np.float32)
    a = ar.astype(np.float32)
    t0 = time.perf_counter()
    if isinstance(rg, sg.SinoPar):
        oracle.back_par(q, np.sin(a), np.cos(a), rg.d, 0.0, xc, yc, ig.mask, dtype=np.float32)
    else:
        oracle.back_fan(q, np.sin(a), np.cos(a), rg.dsd, rg.dso, 0.0, True, rg.d, 0.0, xc, yc,
                        ig.mask, dtype=np.float32)
    tb = time.perf_counter() - t0
    updates = int(ig.mask.sum()) * rg.na

    def err(x, r):
        return float(np.sqrt(((x - r) ** 2).sum() / (r ** 2).sum()))
    return {"rel_rmse_vs_f64_oracle": err(got, ref64), "rel_rmse_vs_f32_oracle": err(got, ref32),
            "cpu_baseline": {"value": updates / tb / 1e9, "unit": "Gupdates/s", "cores": threads,
                             "kind": "port",
                             "sample": f"oracle Float32 backprojection (C++/OpenMP restatement, "
                                       f"not the reference), whole image: {updates} updates in "
                                       f"{tb:.3f} s; CPU {cpu_model()}"}}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 48-reconstruction batch (one HIP graph) leg")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-buffer (PCIe-inclusive) leg")
    ap.add_argument("--no-slices", action="store_true",
                    help="skip the 2-D legs (parallel-beam and arc fan-beam FBP)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned sparse-view correction leg")
    ap.add_argument("--cpu-slices", type=int, default=128)
    ap.add_argument("--view-chunks", type=int, default=None,
                    help="view chunks of a cone-beam execute (the library's default when "
                         "omitted; 1 = one filter launch, then one backprojection)")
    ap.add_argument("--cbct-kernel", choices=("staged", "gather"), default="staged",
                    help="cone-beam backprojector (LDS-staged windows or global gathers)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    dev = torch.device("cuda:0")
    log("generating projections ...")
    sg, rg, ig, ell, proj = make_problem(dev)
    torch.cuda.synchronize()
    updates = int(np.prod(ig.dims)) * rg.na

    log("planning ...")
    ctx = sg._lib.context(0)
    ctx.cbct_kernel = args.cbct_kernel
    if args.view_chunks is not None:
        ctx.view_chunks = args.view_chunks
    chunks = ctx.view_chunks
    forms = ("fp32", "bf16", "f16s")
    plans = {f: sg.plan_fbp(rg, ig, precision=f) for f in forms}
    # Kernel times (filter, backprojection) come from plans that run one
    # filter launch and one backprojection over all views: the chunked
    # execute overlaps them, and its launches each cover a chunk.
    ctx.view_chunks = 1
    serial = {f: sg.plan_fbp(rg, ig, precision=f) for f in forms}
    ctx.view_chunks = chunks
    if args.cbct_kernel == "staged":
        # A/B: the scaled-FP16 form interpolating in packed FP16 arithmetic
        ctx.cbct_kernel = "staged_packed16"
        plans["f16s_pk"] = sg.plan_fbp(rg, ig, precision="f16s")
        serial["f16s_pk"] = plans["f16s_pk"]
        ctx.cbct_kernel = args.cbct_kernel
        forms += ("f16s_pk",)
    outs = {f: sg.recon._empty_colmajor(ig.dims, dev) for f in forms}
    # Untimed warm-up: the GPU's clocks ramp up over the first tens of
    # milliseconds of work (a cold FP32 step measured 3 % slower), so every
    # form runs once and the FP32 plan for about 0.3 s before any timing.
    for f in forms:
        plans[f].execute(proj, out=outs[f])
    t_end = time.perf_counter() + 0.3
    while time.perf_counter() < t_end:
        for _ in range(10):
            plans["fp32"].execute(proj, out=outs["fp32"])
        torch.cuda.synchronize()
    step, kt = {}, {}
    for f in forms:
        log(f"timing {f} ...")
        step[f] = time_steps(lambda: plans[f].execute(proj, out=outs[f]), args.steps,
                             args.warmup)
        kt[f] = kernel_times(serial[f], proj, outs[f], max(3, min(args.steps, 10)))
        log(f"  {f}: step {step[f] * 1e3:.3f} ms (filter {kt[f][0]:.3f} ms, "
            f"back {kt[f][1]:.3f} ms)")
    vols = {f: outs[f].permute(2, 1, 0).contiguous().cpu().numpy().transpose(2, 1, 0)
            for f in forms}
    # PCIe-inclusive rate: host projections in, host volume out (the Julia
    # binding's ccall path), FP32 plan.
    t_host = None
    if not args.no_host:
        log("timing host-buffer executes ...")
        proj_host = np.asfortranarray(proj.permute(2, 1, 0).contiguous().cpu().numpy()
                                      .transpose(2, 1, 0))
        t_host = time_steps(lambda: plans["fp32"].execute(proj_host), 3, 1)

    t32 = step["fp32"]
    f32ms, b32ms = kt["fp32"]
    compulsory = 4 * rg.ns * rg.nt * rg.na + 4 * updates // rg.na
    kname = "cbct_staged_kernel<false,0,false,0>" if args.cbct_kernel == "staged" \
        else "cbct_back_kernel<8,false,0,false>"
    roof = roofline(kname, b32ms, compulsory, updates, "update", 16.0)
    roof["filter_ms"] = f32ms
    ncols = rg.nt * rg.na

    def form(f, bpu):
        fm, bm = kt[f]
        r = {"value": updates / step[f] / 1e9, "unit": "Gupdates/s",
             "ms_per_step": step[f] * 1e3, "speedup_vs_fp32": t32 / step[f],
             "back_kernel_ms": bm, "filter_ms": fm,
             "back_achieved_GBs_model": bpu * updates / (bm * 1e-3) / 1e9,
             "back_speedup_vs_fp32": b32ms / bm}
        r.update(filter_figures(plans[f], fm, rg.ns, ncols, 2))
        return r

    # A/B: the same plans with the Toeplitz (MFMA) filter in place of the FFT form
    toep = {}
    for f in ("fp32", "f16s"):
        log(f"timing {f} with the Toeplitz filter ...")
        pt = plan_toeplitz(sg, rg, ig, precision=f)
        t = time_steps(lambda: pt.execute(proj, out=outs[f]), args.steps, args.warmup)
        ctx.view_chunks = 1
        ps = plan_toeplitz(sg, rg, ig, precision=f)
        ctx.view_chunks = chunks
        fm, bm = kernel_times(ps, proj, outs[f], max(3, min(args.steps, 10)))
        del ps
        toep[f] = {"ms_per_step": t * 1e3, "filter_ms": fm, "back_kernel_ms": bm,
                   "speedup_vs_fp32_toeplitz": None}
        toep[f].update(filter_figures(pt, fm, rg.ns, ncols, 4 if f == "fp32" else 2))
        log(f"  {f} (Toeplitz filter): step {t * 1e3:.3f} ms (filter {fm:.3f} ms)")
        del pt
    toep["f16s"]["speedup_vs_fp32_toeplitz"] = toep["fp32"]["ms_per_step"] / \
        toep["f16s"]["ms_per_step"]
    del toep["fp32"]["speedup_vs_fp32_toeplitz"]
    result = {
        "metric": METRIC,
        "value": updates / t32 / 1e9,
        "unit": "Gupdates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: 12 random ellipsoids (seed 0), analytic Float64 cone-beam "
                "projections cast to Float32",
        "config": {"workload": "FDK CtFanFlat 256x256x128 volume, 512x384 detector, 360 views",
                   "updates_per_step": updates, "projections_per_s": rg.na / t32,
                   "view_chunks": chunks,
                   **filter_figures(plans["fp32"], f32ms, rg.ns, ncols, 4)},
        "roofline": roof,
        "pcie_inclusive": None if t_host is None else {
            "value": updates / t_host / 1e9, "unit": "Gupdates/s", "ms_per_step": t_host * 1e3,
            "note": "FP32 plan, host projections in and host volume out (pageable numpy arrays)"},
        "mixed_f16s": form("f16s", 8.0),
        "mixed_bf16": form("bf16", 8.0),
        "mixed_f16s_packed16": form("f16s_pk", 8.0) if "f16s_pk" in plans else None,
        "toeplitz_filter_ab": toep,
    }
    if not args.no_slices:
        result["slices"] = slice_legs(sg, dev, args.steps, args.warmup, not args.no_cpu)
        log("timing the distance-driven projector ...")
        result["dd_projector"] = dd_leg(sg, dev, max(3, min(args.steps, 10)))
        log("timing the analytic projector ...")
        result["analytic_projector"] = phantom_leg(sg, dev, max(3, min(args.steps, 10)))
        log("timing the filter forms ...")
        result["filter_forms"] = filter_forms_leg(sg, dev, max(3, min(args.steps, 10)))
    if not args.no_learned:
        log("learned correction leg ...")
        result["learned"] = learned_leg(sg, dev, max(3, min(args.steps, 20)))
    if not args.no_batch:
        log("batch leg ...")
        result["batch"] = batch_leg(sg, dev, max(3, min(args.steps, 10)))
    if not args.no_cpu:
        log("CPU oracle leg ...")
        cpu, acc = cpu_leg(sg, rg, ig, proj, plans["fp32"].filter, vols, args.cpu_slices)
        result["cpu_baseline"] = cpu
        result["accuracy"] = acc
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
