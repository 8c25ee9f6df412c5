"""Writes the golden transform vectors tests/golden/sht_T{5,10}.npz.

Independent brute force in 40-digit mpmath: Gauss nodes by Newton on mpmath's own
Legendre polynomial, associated Legendre functions from the exact polynomial derivative
P_n^m = sqrt((2n+1)(n-m)!/(n+m)!) (1-x^2)^{m/2} d^m P_n/dx^m (no Condon-Shortley phase,
int P^2 = 2, space_discretization.tex:160-170), H_n^m = (1-x^2) dP_n^m/dx from the same
exact polynomials, and direct O(L*I*J) sums -- no FFT, no recurrence, no even/odd split.
The reference's own library (SPMODEL/ISPACK) is not vendored, so these vectors pin the
CPU oracle to the documented mathematics.

Run: python tests/golden/gen_golden.py
"""
import os
from fractions import Fraction
from math import comb, factorial

import mpmath as mp
import numpy as np

mp.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))


def legendre_poly(n):
    """Exact coefficients c[k] of P_n(x) = sum c[k] x^k."""
    c = [Fraction(0)] * (n + 1)
    for k in range(n // 2 + 1):
        c[n - 2 * k] += Fraction((-1) ** k * comb(n, k) * comb(2 * n - 2 * k, n), 2 ** n)
    return c


def deriv(c):
    return [c[k] * k for k in range(1, len(c))] or [Fraction(0)]


def peval(c, x):
    acc = mp.mpf(0)
    for k in reversed(range(len(c))):
        acc = acc * x + mp.mpf(c[k].numerator) / c[k].denominator
    return acc


def gauss(J):
    xs, ws = [], []
    for k in range(J):
        x = mp.cos(mp.pi * (k + mp.mpf(3) / 4) / (J + mp.mpf(1) / 2))
        for _ in range(100):
            p = mp.legendre(J, x)
            dp = J * (x * p - mp.legendre(J - 1, x)) / (x * x - 1)
            dx = p / dp
            x -= dx
            if abs(dx) < mp.mpf(10) ** -36:
                break
        dp = J * (x * mp.legendre(J, x) - mp.legendre(J - 1, x)) / (x * x - 1)
        xs.append(x)
        ws.append(2 / ((1 - x * x) * dp * dp))
    order = sorted(range(J), key=lambda i: xs[i])
    return [xs[i] for i in order], [ws[i] for i in order]


def assoc(n, m, x, polys):
    """P_n^m(x), H_n^m(x) = (1-x^2) dP_n^m/dx and cos(phi) dP_n^m/dx = dP_n^m/dphi
    (normalised; x = sin(phi), so cos(phi) = (1-x^2)^{1/2})."""
    Dm = polys[(n, m)]
    Dm1 = polys[(n, m + 1)]
    s2 = 1 - x * x
    norm = mp.sqrt(mp.mpf((2 * n + 1) * factorial(n - m)) / factorial(n + m))
    P = norm * s2 ** (mp.mpf(m) / 2) * peval(Dm, x)
    # d/dx [(1-x^2)^{m/2} D_m] = (1-x^2)^{m/2} D_{m+1} - m x (1-x^2)^{m/2-1} D_m
    H = norm * s2 ** (mp.mpf(m) / 2) * (s2 * peval(Dm1, x) - m * x * peval(Dm, x))
    # the latitude derivative, evaluated with its own powers of (1-x^2) rather than from H
    dphi = norm * (s2 ** (mp.mpf(m + 1) / 2) * peval(Dm1, x) - m * x * s2 ** (mp.mpf(m - 1) / 2) * peval(Dm, x))
    return P, H, dphi


def make(N, I, J, seed):
    rng = np.random.default_rng(seed)
    L = (N + 1) ** 2
    polys = {}
    for n in range(N + 1):
        c = legendre_poly(n)
        for m in range(n + 2):
            polys[(n, m)] = c
            c = deriv(c)
    mu, wl = gauss(J)
    wq = [w / 2 for w in wl]
    lam = [2 * mp.pi * i / I for i in range(I)]
    PT = {}
    HT = {}
    DT = {}
    for n in range(N + 1):
        for m in range(n + 1):
            for j in range(J):
                PT[n, m, j], HT[n, m, j], DT[n, m, j] = assoc(n, m, mu[j], polys)
    # physical longitude and latitude of each node: phi = asin(mu), cos(phi) from phi itself
    coslat = [mp.cos(mp.asin(x)) for x in mu]
    nn = np.floor(np.sqrt(np.arange(L))).astype(int)
    w_in = rng.standard_normal(L) * (nn + 1.0) ** -1.5
    xy_in = rng.standard_normal((J, I))

    def lidx(n, m):
        return n * (n + 1) + m

    def inverse(kind):
        """Direct sums.  gradlon = (1/cos phi) d/dlambda and gradlat = d/dphi: the unit-sphere
        gradient components that the reference's callers divide by the radius
        (HBEBaroc_hspm_vfvm_mod.f90:403,416; LPhys_RediGM_hspm_vfvm_mod.f90:433-434)."""
        table = {"gradmu": HT, "gradlat": DT}.get(kind, PT)
        dlam = kind in ("gradlambda", "gradlon")
        out = np.zeros((J, I))
        for j in range(J):
            for i in range(I):
                acc = mp.mpf(0)
                for n in range(N + 1):
                    if not dlam:
                        acc += mp.mpf(w_in[lidx(n, 0)]) * table[n, 0, j]
                    for m in range(1, n + 1):
                        a, b = mp.mpf(w_in[lidx(n, m)]), mp.mpf(w_in[lidx(n, -m)])
                        T = table[n, m, j]
                        c, s = mp.cos(m * lam[i]), mp.sin(m * lam[i])
                        if dlam:
                            acc += 2 * m * (-a * s - b * c) * T
                        else:
                            acc += 2 * (a * c - b * s) * T
                if kind == "gradlon":
                    acc /= coslat[j]
                out[j, i] = float(acc)
        return out

    def forward(kind):
        out = np.zeros(L)
        G = {}
        for j in range(J):
            for m in range(N + 1):
                re = mp.fsum(mp.mpf(xy_in[j, i]) * mp.cos(m * lam[i]) for i in range(I)) / I
                im = -mp.fsum(mp.mpf(xy_in[j, i]) * mp.sin(m * lam[i]) for i in range(I)) / I
                G[j, m] = (re, im)
        for n in range(N + 1):
            for m in range(n + 1):
                re = mp.mpf(0)
                im = mp.mpf(0)
                for j in range(J):
                    if kind == "val":
                        f = wq[j] * PT[n, m, j]
                    elif kind == "divlambda":
                        f = wq[j] / (1 - mu[j] ** 2) * PT[n, m, j]
                    else:
                        f = -wq[j] / (1 - mu[j] ** 2) * HT[n, m, j]
                    re += f * G[j, m][0]
                    im += f * G[j, m][1]
                if kind == "divlambda":
                    re, im = -m * im, m * re
                out[lidx(n, m)] = float(re)
                if m > 0:
                    out[lidx(n, -m)] = float(im)
        return out

    return dict(
        N=N, I=I, J=J,
        mu=np.array([float(x) for x in mu]), wq=np.array([float(x) for x in wq]),
        w_in=w_in, xy_in=xy_in,
        inv_val=inverse("val"), inv_gradlambda=inverse("gradlambda"), inv_gradmu=inverse("gradmu"),
        inv_gradlon=inverse("gradlon"), inv_gradlat=inverse("gradlat"),
        fwd_val=forward("val"), fwd_divlambda=forward("divlambda"), fwd_divmu=forward("divmu"),
    )


if __name__ == "__main__":
    for (N, I, J, seed) in [(5, 16, 8, 5), (10, 32, 16, 10)]:
        d = make(N, I, J, seed)
        path = os.path.join(HERE, f"sht_T{N}.npz")
        np.savez(path, **d)
        print("wrote", path)
