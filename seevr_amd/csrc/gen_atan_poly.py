"""Fit the odd polynomial used by the library's atan2f (sv_math.h).

Method: after reduction to |z| <= tan(pi/8) (atan t = pi/4 + atan((t-1)/(t+1))
for t > tan(pi/8)), atan z ~= z + z^3 * P(z^2) with P of degree 4.  P is fitted
by iteratively reweighted least squares on the relative error (a discrete
approximation of the minimax fit, Lawson's algorithm), then rounded to float32.
Prints the coefficients as C hex-float literals.

Run: python seevr_amd/csrc/gen_atan_poly.py
"""
import numpy as np

DEG = 5  # number of coefficients of P
LIM = np.tan(np.pi / 8)


def fit():
    z = np.cos(np.linspace(0, np.pi / 2, 20001)) * LIM  # Chebyshev-like spacing
    z = z[z > 1e-6]
    u = z * z
    target = (np.arctan(z) - z) / z ** 3  # P(u) must approximate this
    a = np.stack([u ** k for k in range(DEG)], axis=1)
    w = np.ones_like(z)
    for _ in range(60):  # Lawson reweighting toward equiripple relative error
        sw = np.sqrt(w)
        c, *_ = np.linalg.lstsq(a * sw[:, None], target * sw, rcond=None)
        err = np.abs((z + z ** 3 * (a @ c)) - np.arctan(z)) / np.arctan(z)
        w = w * (err / err.max() + 1e-3)
        w /= w.sum()
    return c


def main():
    c = fit()
    c32 = c.astype(np.float32)
    z = np.linspace(1e-4, LIM, 200001).astype(np.float32)
    u = z * z
    p = np.float32(0)
    for k in reversed(range(DEG)):
        p = (p * u + c32[k]).astype(np.float32)
    r = (z + (z * u) * p).astype(np.float32)
    rel = np.abs(r.astype(np.float64) - np.arctan(z.astype(np.float64))) / np.arctan(z)
    print(f"/* max relative error on (0, tan(pi/8)]: {rel.max():.3e} */")
    for k, v in enumerate(c32):
        print(f"#define SV_ATAN_C{k} {float(v).hex()}f  /* {float(v):.9e} */")


if __name__ == "__main__":
    main()
