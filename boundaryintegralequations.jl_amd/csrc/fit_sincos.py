"""Derives the polynomial coefficients of the device sin/cos kernels (sincos_f64 in
kernels_math.hpp) on |r| <= pi/4 by a discrete minimax (Remez-exchange) fit in
high precision, then rounds them to float64.

  sin(r) ~ r + r^3 * S(r^2),  S of degree DS
  cos(r) ~ 1 - r^2/2 + r^4 * C(r^2), C of degree DC

Run:  python fit_sincos.py   (prints the coefficient tables and their max error)
"""
import mpmath as mp

mp.mp.dps = 60
PI4 = mp.pi / 4


def remez(f, basis, a, b, n_iter=30, ngrid=4000):
    """Minimise max|f(t) - sum c_k basis_k(t)| on [a, b] (t = r^2)."""
    nb = len(basis)
    # initial reference: Chebyshev extrema
    ref = [a + (b - a) * (1 - mp.cos(mp.pi * i / nb)) / 2 for i in range(nb + 1)]
    for _ in range(n_iter):
        A = mp.matrix(nb + 1, nb + 1)
        rhs = mp.matrix(nb + 1, 1)
        for i, t in enumerate(ref):
            for k in range(nb):
                A[i, k] = basis[k](t)
            A[i, nb] = (-1) ** i
            rhs[i] = f(t)
        sol = mp.lu_solve(A, rhs)
        c = [sol[k] for k in range(nb)]
        err = lambda t: f(t) - sum(c[k] * basis[k](t) for k in range(nb))
        grid = [a + (b - a) * mp.mpf(i) / ngrid for i in range(ngrid + 1)]
        vals = [err(t) for t in grid]
        # new reference: alternating extrema
        ext = [0]
        for i in range(1, ngrid):
            if (vals[i] - vals[i - 1]) * (vals[i + 1] - vals[i]) <= 0:
                ext.append(i)
        ext.append(ngrid)
        ext = sorted(set(ext), key=lambda i: i)
        # keep nb+1 points with alternating signs and largest magnitude
        pts = []
        for i in ext:
            if pts and mp.sign(vals[i]) == mp.sign(vals[pts[-1]]):
                if abs(vals[i]) > abs(vals[pts[-1]]):
                    pts[-1] = i
            else:
                pts.append(i)
        while len(pts) > nb + 1:
            if abs(vals[pts[0]]) < abs(vals[pts[-1]]):
                pts.pop(0)
            else:
                pts.pop()
        if len(pts) < nb + 1:
            break
        ref = [grid[i] for i in pts]
    return c, max(abs(v) for v in vals)


def fit(DS, DC):
    b = PI4 ** 2
    # sin: weight so that the absolute error of r^3 S(r^2) is minimised on r in [0, pi/4]
    fs = lambda t: (mp.sin(mp.sqrt(t)) - mp.sqrt(t)) / mp.sqrt(t) ** 3 if t > 0 else mp.mpf(-1) / 6
    fc = lambda t: (mp.cos(mp.sqrt(t)) - 1 + t / 2) / t ** 2 if t > 0 else mp.mpf(1) / 24
    bs = [(lambda k: (lambda t: t ** k))(k) for k in range(DS + 1)]
    bc = [(lambda k: (lambda t: t ** k))(k) for k in range(DC + 1)]
    cs, _ = remez(fs, bs, mp.mpf(0), b)
    cc, _ = remez(fc, bc, mp.mpf(0), b)
    cs = [float(x) for x in cs]
    cc = [float(x) for x in cc]
    # evaluate the float64-rounded polynomials (in high precision) against sin/cos
    worst_s = worst_c = mp.mpf(0)
    for i in range(1, 3001):
        r = PI4 * i / 3000
        t = r * r
        S = sum(mp.mpf(cs[k]) * t ** k for k in range(DS + 1))
        C = sum(mp.mpf(cc[k]) * t ** k for k in range(DC + 1))
        worst_s = max(worst_s, abs(r + r ** 3 * S - mp.sin(r)))
        worst_c = max(worst_c, abs(1 - t / 2 + t * t * C - mp.cos(r)))
    return cs, cc, float(worst_s), float(worst_c)


if __name__ == "__main__":
    for DS, DC in [(4, 4), (5, 5), (5, 6)]:
        cs, cc, es, ec = fit(DS, DC)
        print(f"DS={DS} DC={DC}  max abs err sin {es:.2e} cos {ec:.2e}")
        print("  S =", ", ".join(repr(x) for x in cs))
        print("  C =", ", ".join(repr(x) for x in cc))
