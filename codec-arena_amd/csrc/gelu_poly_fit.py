"""Coefficients of the packed-f16 GELU used by the MLP screen (mlp_common.h).

GELU (tanh form) is x * S(x), S(x) = 0.5 (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))).
The screen evaluates S on packed f16 pairs without transcendentals:

    x = f16(x)                               round toward zero (v_cvt_pkrtz)
    v = x Q,       Q = f16(sqrt 2) / 4
    w = v v - 1                              in [-1, 1] for |x| <= 4.0004
    S = clamp(1/2 + v R(w), 0, 1)            R of degree 6, Horner in w;
                                             the clamp is the FMA's clamp bit

so |x| <= 4 / f16(sqrt 2) * sqrt 2 (about 4.0004) covers the fitted range
w in [-1, 1].  Beyond it w is not clamped: R has even degree and a positive
leading coefficient and stays above 1 / (2 |v|) for every w > 1 (up to f16
overflow, where it is +inf), so v R(w) passes +-1/2 and the clamp saturates S
(the true S is within 1.7e-5 of 0 or 1 there).  The emulation below checks
this out to the f16 range of x.  R is fitted by iteratively reweighted least squares on Chebyshev
nodes in w (a practical route to the minimax fit), then rounded to f16 from
the highest coefficient down, refitting the remaining ones after each
rounding.  Printing the table and the error of an f16 emulation of the
evaluation sequence:

    python gelu_poly_fit.py
"""
import numpy as np

K, C = np.sqrt(2.0 / np.pi), 0.044715
DEG = 6


def s_exact(x):
    return 0.5 * (1.0 + np.tanh(K * (x + C * x ** 3)))


def h(v):
    with np.errstate(over="ignore"):
        return np.float16(v).astype(np.float64)


B = float(np.float16(np.sqrt(2.0)))
Q = B / 4.0          # exact in f16
KX = 1.0 / Q         # x = KX v


def fit():
    w = np.cos(np.pi * (np.arange(8000) + 0.5) / 8000)
    v = np.sqrt(w + 1.0)
    keep = v > 1e-6
    v, w = v[keep], w[keep]
    y = (s_exact(KX * v) - 0.5) / v
    coef = np.zeros(DEG + 1)
    fixed = {}
    for top in range(DEG, -1, -1):
        free = [k for k in range(DEG + 1) if k not in fixed]
        rhs = y - sum(fixed[k] * w ** k for k in fixed)
        M = np.stack([w ** k for k in free], 1)
        wt = v.copy()
        for _ in range(40):
            sol = np.linalg.lstsq(M * wt[:, None], rhs * wt, rcond=None)[0]
            err = np.abs(v * (M @ sol - rhs))
            wt = wt * (1.0 + err / err.max()) ** 2
        for k, c in zip(free, sol):
            coef[k] = c
        fixed[top] = float(np.float16(coef[top]))
    return np.array([fixed[k] for k in range(DEG + 1)])


def rtz16(x):
    y = h(x)
    over = np.abs(y) > np.abs(x)
    y[over] = np.nextafter(np.float16(y[over]), np.float16(0.0)).astype(np.float64)
    return np.clip(y, -65504.0, 65504.0)


def s_f16(x, coef):
    """The device sequence, every operation rounded to f16 (fma: once)."""
    x16 = rtz16(x)
    v = h(x16 * Q)
    w = h(v * v - 1.0)
    p = np.full_like(x, coef[DEG])
    with np.errstate(over="ignore", invalid="ignore"):
        for c in coef[DEG - 1::-1]:
            p = h(p * w + c)
        return np.clip(h(v * p + 0.5), 0.0, 1.0), x16


if __name__ == "__main__":
    coef = fit()
    x = np.concatenate([np.linspace(-16.0, 16.0, 640001), np.linspace(-70000.0, 70000.0, 280001)])
    s16, x16 = s_f16(x, coef)
    g16 = h(x16 * s16)
    ref = x * s_exact(x)
    print("R =", ", ".join(f"{c:.9g}" for c in coef))
    print(f"max |S_f16 - S| = {np.max(np.abs(s16 - s_exact(x))):.3e}")
    near = np.abs(x) <= 16.0
    print(f"max |gelu_f16 - gelu| = {np.max(np.abs(g16 - ref)[near]):.3e} on [-16, 16]")
    big = near & (np.abs(ref) > 0.01)
    print(f"max relative gelu error where |gelu| > 0.01: {np.max(np.abs(g16 - ref)[big] / np.abs(ref[big])):.3e}")
