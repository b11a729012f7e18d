"""Derive the coefficients of the FP32 arctangent used by the 2-D fan-beam
backprojector (csrc/k_back2d.hip, atan2_fast).

atan(u) ~ u * P(u^2) on [0, 1], P of degree 9: minimax in relative error,
found by Lawson's iteratively reweighted least squares in float64 on a dense
grid (Chebyshev-clustered), then rounded to float32.  The script reports the
largest error of the float32 Horner/FMA evaluation against float64 atan in
units in the last place (about 1.3 ulp for these coefficients).

Usage: python gen_atan_coeffs.py
"""

import numpy as np

K = 9


def fit(k, iters=60):
    x = np.cos(np.linspace(0, np.pi / 2, 4000))
    x = np.unique(np.concatenate([x, np.linspace(1e-6, 1, 4000)]))
    y = np.arctan(x) / x
    a = np.stack([x ** (2 * j) for j in range(k + 1)], 1)
    w = np.ones_like(x)
    for _ in range(iters):
        sw = np.sqrt(w)
        c, *_ = np.linalg.lstsq(a * sw[:, None] / y[:, None], sw, rcond=None)
        e = np.abs(a @ c / y - 1)
        w = w * (e + 1e-30)
        w /= w.sum()
    return np.float32(c)


def eval_f32(c, u):
    u = u.astype(np.float32)
    u2 = (u * u).astype(np.float32)
    p = np.full_like(u, c[-1])
    for ck in c[-2::-1]:
        p = (p.astype(np.float64) * u2 + np.float64(ck)).astype(np.float32)   # one rounding (FMA)
    return (u * p).astype(np.float32)


def main():
    c = fit(K)
    u = np.concatenate([np.random.default_rng(0).random(2_000_000),
                        [0.0, 1e-7, 1e-3, 0.5, 1.0]]).astype(np.float32)
    ref = np.arctan(u.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    err = np.abs(eval_f32(c, u).astype(np.float64) - ref) / np.maximum(ulp, 1e-45)
    print("coefficients (u^0 .. u^18):")
    print(",\n".join(f"    {v:.9e}f" for v in c))
    print(f"max error {err.max():.2f} ulp")


if __name__ == "__main__":
    main()
