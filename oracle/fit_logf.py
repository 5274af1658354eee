"""Derivation of the coefficients of the single-precision log used by the
oracle (ora_f32.c: ora_logf) and the device (exa_device.h: exa_logf).

After the reduction x = m * 2^e with m in [sqrt(1/2), sqrt(2)), log(m) =
log1p(f) with f = m - 1 in [-0.2929, 0.4142] is evaluated as f + f^2 Q(f),
Q a polynomial of degree 8 fitted to (log1p(f) - f) / f^2.  The fit weights
the error by |f| (the error of Q enters the result as f^2 dQ against a
result of about f) and reweights towards the worst nodes (Lawson's method)
to approach the minimax polynomial.  Prints the coefficients, rounded to
float32, lowest degree first.  Test infrastructure: nothing here is loaded
by the library.
"""
import math

import numpy as np


def fit(deg=8, nodes=4000, rounds=200):
    lo, hi = math.sqrt(0.5) - 1.0, math.sqrt(2.0) - 1.0
    k = np.arange(nodes)
    f = (lo + hi) / 2 + (hi - lo) / 2 * np.cos(np.pi * (k + 0.5) / nodes)
    f = f[np.abs(f) > 1e-6]
    g = (np.log1p(f) - f) / f ** 2
    w = np.abs(f)
    V = np.vander(f, deg + 1, increasing=True)
    wt = np.ones_like(f)
    c = None
    for _ in range(rounds):
        c, *_ = np.linalg.lstsq(V * (w * np.sqrt(wt))[:, None], g * w * np.sqrt(wt), rcond=None)
        err = np.abs((V @ c - g) * w)
        wt = wt * (err / err.max() + 1e-3)
        wt /= wt.sum() / len(wt)
    return c.astype(np.float32)


if __name__ == "__main__":
    for i, c in enumerate(fit()):
        print(f"c{i} = {float(c)!r}f")
