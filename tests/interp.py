"""Independent point interpolation for receiver tests: element search by barycentric coordinates
and Lagrange weights from a monomial Vandermonde in physical coordinates (numpy)."""
import itertools

import numpy as np


def exponents(dim, p):
    return [e for e in itertools.product(range(p + 1), repeat=dim) if sum(e) <= p]


def interp_weights(mesh, x):
    """(element, weights) with p(x) = sum_n w[n] * p[element*Np + n] for the affine mesh."""
    d, Np = mesh.dim, mesh.Np
    X = mesh.node_xyz.reshape(mesh.K, Np, 3)[:, :, :d]
    x = np.asarray(x, dtype=np.float64)[:d]
    v0 = X[:, 0, :]
    E = X[:, 1:d + 1, :] - v0[:, None, :]           # (K, d edges, d coords)
    lam = np.linalg.solve(np.transpose(E, (0, 2, 1)), (x - v0)[..., None])[..., 0]
    score = np.minimum(1 - lam.sum(axis=1), lam.min(axis=1))
    e = int(np.argmax(score))
    assert score[e] > -1e-9, "point outside the mesh"
    c = X[e].mean(axis=0)
    h = np.abs(X[e] - c).max()
    ex = np.array(exponents(d, mesh.order))
    V = np.prod(((X[e] - c) / h)[:, None, :] ** ex[None, :, :], axis=2)   # (Np nodes, Np monomials)
    mono = np.prod(((x - c) / h)[None, :] ** ex, axis=1)
    return e, np.linalg.solve(V.T, mono)


def sample(mesh, u, points):
    """Pressure of state u (4 x K*Np) at the points."""
    out = []
    for x in points:
        e, w = interp_weights(mesh, x)
        out.append(float(w @ u[0][e * mesh.Np:(e + 1) * mesh.Np]))
    return np.array(out)
