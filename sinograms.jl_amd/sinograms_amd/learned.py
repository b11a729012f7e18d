"""Learned sparse-view / short-scan correction for fan-beam FBP.

No counterpart in the reference.  A small 1-D convolutional network f
(width-9 convolutions 1 -> C -> ... -> C -> 1 channels, GELU between them,
BF16-rounded weights, FP32 accumulation; csrc/k_nn.hip) corrects every view
of the filtered sinogram before backprojection: q' = q + f(q).  It starts at
f = 0, i.e. at plain FBP, which stays the fallback and the yardstick.

Training (all device work in libsgx): phantoms are projected by the
distance-driven projector (project_bdd, K6) at the sparse view set and at a
dense full-scan view set; the target is the plain FBP of the dense
projections.  One step over a batch of B phantoms:
  q   = weights + ramp filter of the sparse sinograms       (sgx_plan_filter)
  q'  = q + f(q)                                             (sgx_net_forward)
  x   = backprojection of q'                                 (sgx_plan_back)
  L   = mean (x - target)^2, g = dL/dx                       (sgx_mse)
  dq' = exact transpose of the backprojection applied to g   (sgx_plan_back_t)
  dL/dparams by the network's backward kernels, then Adam    (sgx_net_backward, _adam)
"""

from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import _lib, phantom
from .geom import ImageGeom, SinoFanFlat
from .recon import _empty_colmajor, plan_fbp
from .sys2 import project_bdd


def _cm(t):
    """Column-major float32 copy of a 2-D CUDA tensor."""
    import torch
    return t.to(torch.float32).t().contiguous().t()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class ConvCorrection:
    """The correction network's parameters and kernels on one device."""

    def __init__(self, layers: int = 4, channels: int = 32, seed: int = 0, device: int = 0):
        self._ctx = _lib.context(device)
        self.layers, self.channels, self.device = layers, channels, device
        h = ctypes.c_void_p()
        self._ctx.check(_lib.load().sgx_net_create(self._ctx.h, layers, channels, seed,
                                                   ctypes.byref(h)))
        self.h = h

    @property
    def num_params(self) -> int:
        return _lib.load().sgx_net_num_params(self.h)

    def params(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        self._ctx.check(_lib.load().sgx_net_params(self.h, out.ctypes.data_as(ctypes.c_void_p), 0))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.num_params:
            raise ValueError("parameter count")
        self._ctx.check(_lib.load().sgx_net_params(self.h, p.ctypes.data_as(ctypes.c_void_p), 1))

    def grads(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        self._ctx.check(_lib.load().sgx_net_grads(self.h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def zero_grad(self) -> None:
        self._ctx.check(_lib.load().sgx_net_zero_grad(self.h))

    def forward(self, q, out=None, keep: bool = False):
        """q: (n, rows) column-major float32 CUDA tensor (each column one view
        row); returns q + f(q) in the same layout."""
        n, rows = q.shape
        if out is None:
            out = _empty_colmajor((n, rows), q.device)
        with self._ctx.ordered():
            self._ctx.check(_lib.load().sgx_net_forward(self.h, _ptr(q), rows, n, _ptr(out),
                                                        int(keep)))
        return out

    def backward(self, gout) -> None:
        with self._ctx.ordered():
            self._ctx.check(_lib.load().sgx_net_backward(self.h, _ptr(gout)))

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        self._ctx.check(_lib.load().sgx_net_adam(self.h, lr, beta1, beta2, eps))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.load().sgx_net_destroy(self.h)
                self.h = None
        except Exception:
            pass


def mse(ctx, x, t, grad=None, scale: float = 1.0) -> float:
    """sum (x - t)^2 of two CUDA tensors; grad (optional) = scale * 2 (x - t)."""
    s = ctypes.c_double()
    with ctx.ordered():
        ctx.check(_lib.load().sgx_mse(ctx.h, _ptr(x), _ptr(t), x.numel(), scale,
                                      _ptr(grad) if grad is not None else None, ctypes.byref(s)))
    return s.value


def psnr(x, ref) -> float:
    """PSNR of x against ref, peak = max |ref|."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    m = np.mean((x - ref) ** 2)
    return float(10 * math.log10(np.abs(ref).max() ** 2 / m))


def ssim(x, ref) -> float:
    """Mean SSIM (Gaussian window sigma 1.5, K1 0.01, K2 0.03, range of ref)."""
    from scipy.ndimage import gaussian_filter
    x = np.asarray(x, np.float64)
    y = np.asarray(ref, np.float64)
    L = y.max() - y.min()
    c1, c2 = (0.01 * L) ** 2, (0.03 * L) ** 2
    f = lambda a: gaussian_filter(a, 1.5)
    mx, my = f(x), f(y)
    sxx, syy, sxy = f(x * x) - mx * mx, f(y * y) - my * my, f(x * y) - mx * my
    s = ((2 * mx * my + c1) * (2 * sxy + c2)) / ((mx * mx + my * my + c1) * (sxx + syy + c2))
    return float(s.mean())


class SparseViewProblem:
    """Flat-detector fan-beam geometry (the distance-driven projector models a
    flat detector), a sparse short scan and a dense full scan of the same
    detector, their FBP plans, and phantom sampling through project_bdd."""

    def __init__(self, npix=512, pix=0.98, nb=888, d=1.110, dsd=949.0, dso=541.0,
                 na_sparse=90, na_dense=984, device=0):
        import torch
        self.device = torch.device(f"cuda:{device}")
        kw = dict(nb=nb, d=d, dsd=dsd, dod=dsd - dso)
        short = SinoFanFlat(na=na_dense, **kw).orbit_short()
        self.rg_sparse = SinoFanFlat(na=na_sparse, orbit=short, **kw)
        self.rg_dense = SinoFanFlat(na=na_dense, orbit=360.0, **kw)
        self.ig = ImageGeom((npix, npix), (pix, pix))
        self.plan_sparse = plan_fbp(self.rg_sparse, self.ig, device=device)
        self.plan_dense = plan_fbp(self.rg_dense, self.ig, device=device)
        self.fov = npix * pix
        self.npix, self.pix, self.nb, self.d, self.dsd, self.dso = npix, pix, nb, d, dsd, dso
        self._ctx = _lib.context(device)

    def _geo(self, rg):
        return dict(DSD=self.dsd, DS0=self.dso, pSize=self.pix, dSize=self.d, nPix=self.npix,
                    nDet=self.nb, angle=np.asarray(rg.ar(), np.float64))

    def project(self, img, rg):
        """project_bdd of an image in this package's layout (x first), as the
        (nb, na) column-major float32 sinogram of `rg`."""
        # the DD projector takes rows = y from the top, columns = x
        ph = img.t().flip(0).t().contiguous().t()          # column-major (row, col)
        s = project_bdd(ph, self._geo(rg))                    # (na, nb) column-major, FP64
        return _cm(s.t())

    def sample(self, seed: int):
        """(sparse sinogram, dense-FBP target) of random-ellipse phantom `seed`."""
        import torch
        ell = phantom.random_ellipses(seed, 15, self.fov)
        x, y = self.ig.axes()
        img = phantom.image_ellipses(ell, x, y, device=self.device)
        sparse = self.project(img, self.rg_sparse)
        target = self.plan_dense.execute(self.project(img, self.rg_dense))
        torch.cuda.synchronize()
        return sparse, target

    def filter(self, sino, q=None):
        if q is None:
            q = _empty_colmajor(tuple(sino.shape), sino.device)
        with self._ctx.ordered():
            self._ctx.check(_lib.load().sgx_plan_filter(self.plan_sparse.h, _ptr(sino), _ptr(q),
                                                        0))
        return q

    def back(self, q, out=None):
        if out is None:
            out = _empty_colmajor(self.ig.dims, q.device)
        with self._ctx.ordered():
            self._ctx.check(_lib.load().sgx_plan_back(self.plan_sparse.h, _ptr(q), _ptr(out), 0))
        return out

    def back_t(self, img, q=None):
        if q is None:
            q = _empty_colmajor(self.rg_sparse.dims(), img.device)
        with self._ctx.ordered():
            self._ctx.check(_lib.load().sgx_plan_back_t(self.plan_sparse.h, _ptr(img), _ptr(q),
                                                        0))
        return q

    def reconstruct(self, sino, net=None):
        """Sparse-view FBP, with the learned correction when `net` is given."""
        q = self.filter(sino)
        if net is not None:
            q = net.forward(q)
        return self.back(q)


def train(problem: SparseViewProblem, net: ConvCorrection, seeds, batch: int = 8,
          lr: float = 1e-3, epochs: int = 1, log=None):
    """Train `net` on the phantoms of `seeds` (sampled once, kept on the
    device); returns per-step losses and timings."""
    import torch
    t0 = time.perf_counter()
    data = [problem.sample(int(s)) for s in seeds]
    t_data = time.perf_counter() - t0
    nb, na = problem.rg_sparse.dims()
    npix = problem.npix * problem.npix
    Q = _empty_colmajor((nb, na * batch), problem.device)
    G = _empty_colmajor((nb, na * batch), problem.device)
    X = _empty_colmajor(problem.ig.dims, problem.device)
    gX = _empty_colmajor(problem.ig.dims, problem.device)
    losses = []
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    steps = 0
    for ep in range(epochs):
        order = np.random.default_rng(ep).permutation(len(data))
        for b0 in range(0, len(order) - batch + 1, batch):
            idx = order[b0:b0 + batch]
            for j, i in enumerate(idx):
                problem.filter(data[i][0], Q[:, j * na:(j + 1) * na])
            out = net.forward(Q, keep=True)
            loss = 0.0
            for j, i in enumerate(idx):
                problem.back(out[:, j * na:(j + 1) * na], X)
                loss += mse(problem._ctx, X, data[i][1], gX, scale=1.0 / (batch * npix))
                problem.back_t(gX, G[:, j * na:(j + 1) * na])
            net.zero_grad()
            net.backward(G)
            net.adam(lr)
            losses.append(loss / (batch * npix))
            steps += 1
            if log and steps % 20 == 0:
                log(f"  step {steps}: loss {losses[-1]:.4g}")
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    return {"losses": losses, "steps": steps, "data_s": t_data, "train_s": t_train,
            "phantoms_per_s": steps * batch / t_train if t_train > 0 else 0.0}


def evaluate(problem: SparseViewProblem, net: ConvCorrection, seeds):
    """PSNR / SSIM of plain and learned sparse-view FBP against the dense FBP
    target, over the phantoms of `seeds`."""
    res = {"psnr_plain": [], "psnr_learned": [], "ssim_plain": [], "ssim_learned": []}
    for s in seeds:
        sino, target = problem.sample(int(s))
        t = target.cpu().numpy()
        plain = problem.reconstruct(sino).cpu().numpy()
        learned = problem.reconstruct(sino, net).cpu().numpy()
        res["psnr_plain"].append(psnr(plain, t))
        res["psnr_learned"].append(psnr(learned, t))
        res["ssim_plain"].append(ssim(plain, t))
        res["ssim_learned"].append(ssim(learned, t))
    return {k: float(np.mean(v)) for k, v in res.items()}
