"""Write tests/golden/basis_mp.json: 1D Gauss-Legendre and Gauss-Lobatto rules
on [0,1] and the GLL Lagrange basis B, G at the GL points, all at 40 digits
with mpmath, for p = 1..8 and Q in {p+1, p+2}.

These are closed-form mathematical objects (roots of P_n and of P'_{n-1},
Lagrange interpolation), computed here independently of both the oracle and
the product, so they pin fem/intrules.cpp:620-830 and
fem/fe/fe_base.cpp:1811-2010 (Poly_1D basis) without running MFEM.

    python tests/golden/make_golden.py
"""
import json
import os

import mpmath as mp

mp.mp.dps = 40


def legendre_coeffs(n):
    """Exact coefficients (lowest degree first) of P_n by the recurrence."""
    from fractions import Fraction
    p0, p1 = [Fraction(1)], [Fraction(0), Fraction(1)]
    if n == 0:
        return p0
    for j in range(2, n + 1):
        a = [Fraction(0)] + [(2 * j - 1) * c for c in p1]
        b = [(j - 1) * c for c in p0] + [Fraction(0), Fraction(0)]
        p0, p1 = p1, [(x - y) / j for x, y in zip(a, b)]
    return p1


def roots(coeffs):
    """Real roots of a polynomial (lowest degree first), sorted."""
    c = [mp.mpf(x.numerator) / x.denominator for x in reversed(coeffs)]
    r = mp.polyroots(c, maxsteps=400, extraprec=400)
    return sorted(mp.re(x) for x in r)


def gauss_legendre(n):
    z = roots(legendre_coeffs(n))
    pc = legendre_coeffs(n)
    dpc = [k * pc[k] for k in range(1, len(pc))]
    xs, ws = [], []
    for r in z:
        dp = mp.polyval([mp.mpf(x.numerator) / x.denominator for x in reversed(dpc)], r)
        xs.append((1 + r) / 2)
        ws.append(1 / ((1 - r * r) * dp * dp))
    return xs, ws


def gauss_lobatto(n):
    N = n - 1
    if n == 1:
        return [mp.mpf(1) / 2], [mp.mpf(1)]
    pc = legendre_coeffs(N)
    dpc = [k * pc[k] for k in range(1, len(pc))]
    xs = [mp.mpf(-1)] + (roots(dpc) if N >= 2 else []) + [mp.mpf(1)]
    pcv = [mp.mpf(x.numerator) / x.denominator for x in reversed(pc)]
    ws = [2 / (N * (N + 1) * mp.polyval(pcv, x) ** 2) for x in xs]
    return [(x + 1) / 2 for x in xs], [w / 2 for w in ws]


def lagrange(nodes, x):
    B, G = [], []
    for d, zd in enumerate(nodes):
        num, den, dnum = mp.mpf(1), mp.mpf(1), mp.mpf(0)
        for k, zk in enumerate(nodes):
            if k == d:
                continue
            dnum = dnum * (x - zk) + num
            num *= (x - zk)
            den *= (zd - zk)
        B.append(num / den)
        G.append(dnum / den)
    return B, G


out = {"gl": {}, "gll": {}, "basis": {}}
for n in range(1, 12):
    x, w = gauss_legendre(n)
    out["gl"][n] = {"x": [mp.nstr(v, 25) for v in x], "w": [mp.nstr(v, 25) for v in w]}
for n in range(2, 11):
    x, w = gauss_lobatto(n)
    out["gll"][n] = {"x": [mp.nstr(v, 25) for v in x], "w": [mp.nstr(v, 25) for v in w]}
for p in range(1, 9):
    nodes, _ = gauss_lobatto(p + 1)
    for Q in (p + 1, p + 2):
        xq, _ = gauss_legendre(Q)
        Bm, Gm = [], []
        for xv in xq:
            b, g = lagrange(nodes, xv)
            Bm.append([mp.nstr(v, 25) for v in b])
            Gm.append([mp.nstr(v, 25) for v in g])
        out["basis"][f"{p},{Q}"] = {"B": Bm, "G": Gm}

path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis_mp.json")
with open(path, "w") as f:
    json.dump(out, f, indent=0)
print("wrote", path)
