"""Derivation of the exp polynomial in csrc/kinc_math.h (kinc_expf).

After the Cody-Waite reduction x = g + n*ln2, |g| <= 0.3467, e^g is
evaluated as 1 + g (1 + g P(g)) with P of degree 4.  P comes from a Remez
exchange (E. Ya. Remez, 1934) on f(g) = (e^g - 1 - g) / g^2 with weight
g^2 / e^g, so that the equioscillating error is the relative error of e^g.

    python kinc_amd/exp_remez.py     # prints the coefficients (float32) and the error

Needs mpmath; nothing in the library imports this file.
"""
from mpmath import cos, exp, lu_solve, matrix, mp, mpf, pi

mp.dps = 40
H = mpf("0.3467")
DEG = 4


def f(g):
    return (exp(g) - 1 - g) / (g * g)


def weight(g):
    return g * g / exp(g)


def remez(n=DEG + 1, iters=30, grid=4000):
    pts = [H * cos(pi * i / n) for i in range(n + 1)]
    for _ in range(iters):
        A = matrix(n + 1, n + 1)
        b = matrix(n + 1, 1)
        for i, x in enumerate(pts):
            for j in range(n):
                A[i, j] = x ** j
            w = weight(x) if abs(x) > mpf("1e-6") else mpf("1e-12")
            A[i, n] = -((-1) ** i) / w
            b[i] = f(x)
        sol = lu_solve(A, b)
        c = [sol[j] for j in range(n)]

        def err(x):
            return (sum(c[j] * x ** j for j in range(n)) - f(x)) * weight(x)

        xs = [-H + 2 * H * k / grid for k in range(grid + 1)]
        es = [err(x) if x != 0 else mpf(0) for x in xs]
        ext = []
        for k, e in enumerate(es):
            left = es[k - 1] if k > 0 else None
            right = es[k + 1] if k < grid else None
            if (left is None or abs(e) >= abs(left)) and (right is None or abs(e) >= abs(right)) \
                    and e != 0:
                ext.append((xs[k], e))
        merged = []
        for x, e in ext:
            if merged and (merged[-1][1] > 0) == (e > 0):
                if abs(e) > abs(merged[-1][1]):
                    merged[-1] = (x, e)
            else:
                merged.append((x, e))
        if len(merged) >= n + 1:
            s = max(range(len(merged) - n),
                    key=lambda s0: min(abs(m[1]) for m in merged[s0:s0 + n + 1]))
            pts = [m[0] for m in merged[s:s + n + 1]]
        worst = max(abs(e) for e in es)
    return c, worst


if __name__ == "__main__":
    import numpy as np
    coef, worst = remez()
    print("max relative error of e^g before rounding: %s" % mp.nstr(worst, 5))
    for j, v in enumerate(coef):
        print("g^%d: %r" % (j, float(np.float32(float(v)))))
