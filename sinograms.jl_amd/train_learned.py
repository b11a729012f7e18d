"""Train and evaluate the learned sparse-view correction; print a JSON report.

Usage (one GPU):
    python sinograms.jl_amd/train_learned.py [--train N] [--eval M] [--epochs E]
        [--npix 512] [--batch 8] [--lr 1e-3] [--out report.json] [--save params.npy]

Geometry: flat-detector fan beam (the distance-driven projector models a flat
detector) with the angular coverage of the 888-bin, 1.0239 mm arc detector at
dso 541 / dsd 949 (888 bins of 1.110 mm), a 512 x 512 image of 0.98 mm; input
90 views over the short scan (180 degrees + fan angle, Parker weights),
target the plain FBP of 984 views over 360 degrees.  Network: 4 layers, 32
channels, width 9, GELU, BF16-rounded weights, FP32 accumulation.  Phantoms:
15 random ellipses per seed (training seeds 0..N-1, evaluation seeds from
10^6), projected by project_bdd.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--eval", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--npix", type=int, default=512)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--chunk", type=int, default=400, help="phantoms resident at a time")
    ap.add_argument("--out", default=None)
    ap.add_argument("--save", default=None)
    args = ap.parse_args()

    import numpy as np
    import torch
    from sinograms_amd import learned

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    pix = 0.98 * 512 / args.npix
    prob = learned.SparseViewProblem(npix=args.npix, pix=pix)
    net = learned.ConvCorrection(layers=4, channels=32, seed=0)
    log(f"network: {net.num_params} parameters; sparse views {prob.rg_sparse.na} over "
        f"{prob.rg_sparse.orbit:.2f} deg, dense {prob.rg_dense.na}")

    # sampling cost
    t0 = time.perf_counter()
    for s in range(4):
        prob.sample(10**7 + s)
    sample_s = (time.perf_counter() - t0) / 4
    log(f"sampling: {sample_s * 1e3:.1f} ms per phantom (two DD projections + dense FBP)")

    # training, in resident chunks of phantoms
    hist = {"losses": [], "steps": 0, "data_s": 0.0, "train_s": 0.0}
    for ep in range(args.epochs):
        for c0 in range(0, args.train, args.chunk):
            seeds = range(c0, min(args.train, c0 + args.chunk))
            h = learned.train(prob, net, seeds, batch=args.batch, lr=args.lr, epochs=1,
                              log=None)
            for k in ("losses",):
                hist[k] += h[k]
            hist["steps"] += h["steps"]
            hist["data_s"] += h["data_s"]
            hist["train_s"] += h["train_s"]
            log(f"epoch {ep} phantoms {c0}..{c0 + len(seeds) - 1}: loss "
                f"{np.mean(h['losses'][-5:]):.4g} ({h['train_s']:.1f} s train, "
                f"{h['data_s']:.1f} s data)")
    # inference throughput (filter + correction + backprojection of one sinogram)
    sino, _ = prob.sample(10**7 + 99)
    for _ in range(3):
        prob.reconstruct(sino, net)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 20
    for _ in range(reps):
        prob.reconstruct(sino, net)
    torch.cuda.synchronize()
    infer_ms = (time.perf_counter() - t0) / reps * 1e3
    t0 = time.perf_counter()
    for _ in range(reps):
        prob.reconstruct(sino)
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) / reps * 1e3

    ev = learned.evaluate(prob, net, range(10**6, 10**6 + args.eval))
    rep = {
        "geometry": "flat fan 888 x 1.110 mm, dso 541, dsd 949; image "
                    f"{args.npix}^2 at {pix:.3f} mm; 90 short-scan views -> 984-view FBP target",
        "network": "4 layers, 32 channels, width 9, GELU, BF16 weights, FP32 accumulation",
        "train_phantoms": args.train, "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
        "eval_phantoms": args.eval, "steps": hist["steps"],
        "loss_first": float(np.mean(hist["losses"][:5])),
        "loss_last": float(np.mean(hist["losses"][-5:])),
        "train_s": hist["train_s"], "data_s": hist["data_s"],
        "train_phantoms_per_s": hist["steps"] * args.batch / max(hist["train_s"], 1e-9),
        "sample_ms_per_phantom": sample_s * 1e3,
        "inference_ms_per_image": infer_ms, "plain_fbp_ms_per_image": plain_ms,
        **ev,
    }
    if args.save:
        np.save(args.save, net.params())
    js = json.dumps(rep)
    print(js, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(js + "\n")


if __name__ == "__main__":
    main()
