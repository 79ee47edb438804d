"""Writes tests/golden/reference_matrices.json: exact reference-element matrices.

Independent of both the product and the oracle: node sets are the Gmsh orderings
published in the Gmsh manual ("Node ordering") for low orders, the Lagrange basis is
obtained by exact rational inversion of a monomial Vandermonde, and integrals use the
closed form  int_T x^a y^b z^c = a! b! c! / (a+b+c+d)!  on the unit simplex of dimension
d (line: int_{-1}^{1} u^a du). Output: M (mass), S[r] (S[r][i][j] = int d_r phi_i phi_j)
and the face mass matrix Mf, as floats (exactly rounded from the rationals).
Run: python tests/golden/make_reference_matrices.py
"""
import itertools
import json
import os
from fractions import Fraction as Fr
from math import factorial

import sympy as sp

u, v, w = sp.symbols("u v w")
X = (u, v, w)

R = Fr


def nodes(dim, p):
    """Gmsh-order reference nodes (manual, 'Node ordering') for the orders fixed here."""
    if dim == 1:
        pts = [(-1,), (1,)] + [(R(-1) + R(2 * i, p),) for i in range(1, p)]
        return pts
    if dim == 2:
        V = [(R(0), R(0)), (R(1), R(0)), (R(0), R(1))]
        pts = list(V)
        for a, b in ((0, 1), (1, 2), (2, 0)):
            for i in range(1, p):
                pts.append(tuple(V[a][c] + (V[b][c] - V[a][c]) * R(i, p) for c in range(2)))
        if p == 3:
            pts.append((R(1, 3), R(1, 3)))
        assert p <= 3
        return pts
    V = [(R(0), R(0), R(0)), (R(1), R(0), R(0)), (R(0), R(1), R(0)), (R(0), R(0), R(1))]
    pts = list(V)
    for a, b in ((0, 1), (1, 2), (2, 0), (3, 0), (3, 2), (3, 1)):
        for i in range(1, p):
            pts.append(tuple(V[a][c] + (V[b][c] - V[a][c]) * R(i, p) for c in range(3)))
    if p == 3:  # one node per face at its centroid, faces (0,2,1), (0,1,3), (0,3,2), (3,1,2)
        for fa in ((0, 2, 1), (0, 1, 3), (0, 3, 2), (3, 1, 2)):
            pts.append(tuple(sum(V[k][c] for k in fa) / 3 for c in range(3)))
    assert p <= 3
    return pts


def monomials(dim, p):
    return [e for e in itertools.product(range(p + 1), repeat=dim) if sum(e) <= p]


def integrate_monomial(dim, e):
    if dim == 1:
        a = e[0]
        return R(0) if a % 2 else R(2, a + 1)
    num = 1
    for k in e:
        num *= factorial(k)
    return R(num, factorial(sum(e) + dim))


def basis(dim, p):
    pts = nodes(dim, p)
    mons = monomials(dim, p)
    n = len(pts)
    V = sp.Matrix(n, n, lambda i, j: sp.prod([sp.Rational(pts[i][c]) ** mons[j][c] for c in range(dim)]))
    Vi = V.inv()
    mexpr = [sp.prod([X[c] ** m[c] for c in range(dim)]) for m in mons]
    phi = [sp.expand(sum(Vi[j, i] * mexpr[j] for j in range(n))) for i in range(n)]
    return pts, phi


def integrate(dim, expr):
    poly = sp.Poly(sp.expand(expr), *X[:dim])
    tot = R(0)
    for mon, coef in poly.terms():
        c = sp.Rational(coef)
        tot += R(int(c.p), int(c.q)) * integrate_monomial(dim, mon)
    return tot


def matrices(dim, p):
    pts, phi = basis(dim, p)
    n = len(phi)
    M = [[integrate(dim, phi[i] * phi[j]) for j in range(n)] for i in range(n)]
    S = [[[integrate(dim, sp.diff(phi[i], X[r]) * phi[j]) for j in range(n)] for i in range(n)] for r in range(dim)]
    return pts, M, S


def f(x):
    return float(x)


def main():
    out = {}
    cases = [(1, 1), (1, 2), (1, 3), (1, 4), (2, 1), (2, 2), (2, 3), (3, 1), (3, 2), (3, 3)]
    for dim, p in cases:
        pts, M, S = matrices(dim, p)
        entry = {"nodes": [[f(c) for c in q] + [0.0] * (3 - dim) for q in pts],
                 "M": [[f(x) for x in row] for row in M],
                 "S": [[[f(x) for x in row] for row in Sr] for Sr in S],
                 "M_rational": [[str(x) for x in row] for row in M]}
        if dim > 1:
            _, Mf, _ = matrices(dim - 1, p)
            entry["Mf"] = [[f(x) for x in row] for row in Mf]
        else:
            entry["Mf"] = [[1.0]]
        out[f"{dim}_{p}"] = entry
        print("done", dim, p, flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_matrices.json")
    with open(path, "w") as fh:
        json.dump(out, fh)
    print("wrote", path)


if __name__ == "__main__":
    main()
