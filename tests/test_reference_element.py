"""Reference-element matrices of the product and of the oracle against exact rationals.

The golden fixture (tests/golden/reference_matrices.json) is computed with exact rational
arithmetic from the Gmsh manual's node orderings (tests/golden/make_reference_matrices.py),
so a wrong node order, a dropped quadrature term or a transposed operand fails here.
"""
import json
import os

import numpy as np
import pytest

import dgk
import oracle_binding as ob

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_matrices.json")))


@pytest.mark.parametrize("key", sorted(GOLD))
def test_product_matches_exact(key):
    dim, p = map(int, key.split("_"))
    g = GOLD[key]
    r = dgk.ref_element(dim, p)
    np.testing.assert_allclose(r["nodes"], np.array(g["nodes"]), atol=1e-15)
    np.testing.assert_allclose(r["M"], np.array(g["M"]), rtol=0, atol=2e-16 * np.abs(g["M"]).max() * 8)
    np.testing.assert_allclose(r["S"], np.array(g["S"]), rtol=0, atol=1e-15 * max(1, np.abs(g["S"]).max()))
    np.testing.assert_allclose(r["Mf"], np.array(g["Mf"]), rtol=0, atol=1e-15)


@pytest.mark.parametrize("key", sorted(GOLD))
def test_oracle_matches_exact(key):
    dim, p = map(int, key.split("_"))
    g = GOLD[key]
    M, S, Mf = ob.ref_matrices(dim, p)
    np.testing.assert_allclose(ob.ref_nodes(dim, p), np.array(g["nodes"]), atol=1e-15)
    np.testing.assert_allclose(M, np.array(g["M"]), rtol=0, atol=2e-15 * np.abs(g["M"]).max())
    np.testing.assert_allclose(S, np.array(g["S"]), rtol=0, atol=2e-15 * max(1, np.abs(g["S"]).max()))
    np.testing.assert_allclose(Mf, np.array(g["Mf"]), rtol=0, atol=2e-15)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_product_vs_oracle_and_invariants(dim, p):
    """All orders: product and oracle agree; measure, partition of unity, SPD mass, exact derivative."""
    r = dgk.ref_element(dim, p)
    M, S, Mf = ob.ref_matrices(dim, p)
    np.testing.assert_allclose(r["nodes"], ob.ref_nodes(dim, p), atol=1e-14)
    np.testing.assert_allclose(r["M"], M, atol=1e-15)
    np.testing.assert_allclose(r["S"], S, atol=1e-14)
    np.testing.assert_allclose(r["Mf"], Mf, atol=1e-15)
    vol = {1: 2.0, 2: 0.5, 3: 1.0 / 6}[dim]
    assert abs(r["M"].sum() - vol) < 1e-14
    assert np.abs(r["S"].sum(axis=1)).max() < 1e-13  # sum_i d phi_i = 0
    assert np.linalg.eigvalsh(r["M"]).min() > 0
    # W_r applied to nodal values of a degree-p polynomial returns the L2 projection of its derivative
    # tested on f = x_r: sum_j W_r[i][j] * 1 ... use weak form: (W_r f)_i = M^-1 int d_r phi_i f
    x = r["nodes"]
    for rr in range(dim):
        f = x[:, rr] ** min(p, 2) + 0.3 * x[:, 0] * (p >= 2)
        lhs = r["M"] @ (r["W"][rr] @ f)
        rhs = r["S"][rr] @ f
        np.testing.assert_allclose(lhs, rhs, atol=1e-13)
    # LIFT = M^-1 E Mf
    Nf, Nfp = r["fmask"].shape
    E = np.zeros((x.shape[0], Nf * Nfp))
    for j in range(Nf):
        for k in range(Nfp):
            E[r["fmask"][j, k], j * Nfp + k] = 1
    L = np.linalg.solve(r["M"], E @ np.kron(np.eye(Nf), r["Mf"]))
    np.testing.assert_allclose(r["LIFT"], L, atol=1e-11 * np.abs(L).max())


def test_p1_closed_forms():
    r = dgk.ref_element(2, 1)
    np.testing.assert_allclose(r["M"] * 24, np.ones((3, 3)) + np.eye(3), atol=1e-14)
    r = dgk.ref_element(3, 1)
    np.testing.assert_allclose(r["M"] * 120, np.ones((4, 4)) + np.eye(4), atol=1e-13)
    # faces of the unit tetrahedron in Gmsh order
    assert r["fmask"].tolist() == [[0, 2, 1], [0, 1, 3], [0, 3, 2], [3, 1, 2]]
