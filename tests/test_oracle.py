"""Pins the CPU oracle before it is trusted as the checker.

Each check is chosen so that a plausible mistake (dropped term, wrong sign or index,
transposed operand) fails it: the two independent evaluation paths of the oracle
(literal quadrature loops with per-element LU inverses vs factorized reference matrices),
the reference's own quadratic face search vs the hashed one, RK4 against the exact
matrix polynomial of the assembled operator, conservation of pressure mass and the energy
bound with reflecting walls, and closed forms for constant states.
"""
import numpy as np
import pytest

import dgk
import oracle_binding as ob
from workloads import conf


def rect(p, nx=4, ny=3, bcs="Wall=Reflecting\nAbs=Absorbing\n", side=(1, 2, 1, 2)):
    return dgk.gen_rect(nx, ny, 1.0, 0.75, p, list(side), "1:Wall;2:Abs", conf(bcs))


def box(p, n=2, bcs="Wall=Reflecting\nAbs=Absorbing\n", side=(1, 1, 2, 2, 1, 2), jitter=0.1):
    return dgk.gen_box(n, n, n, 1.0, 0.9, 0.8, p, jitter, 2, list(side), "1:Wall;2:Abs", conf(bcs))


@pytest.mark.parametrize("dim,p", [(1, 1), (1, 3), (2, 1), (2, 2), (2, 3), (2, 4), (3, 1), (3, 2), (3, 3), (3, 4)])
@pytest.mark.parametrize("sign", [ob.SIGN_REFERENCE, ob.SIGN_STABLE])
def test_literal_equals_factorized(dim, p, sign):
    if dim == 1:
        m = dgk.gen_line(7, 1.0, p, [1, 2], "1:Wall;2:Abs", conf("Wall=Reflecting\nAbs=Absorbing\n"))
    else:
        m = rect(p) if dim == 2 else box(p)
    v0 = (40.0, -20.0, 10.0 if dim == 3 else 0.0)
    o = ob.Oracle(m, 340.0, 1.225, v0=v0, sign_mode=sign)
    u = np.random.default_rng(7).standard_normal((4, m.K * m.Np))
    if dim < 3:
        u[3] = 0
    a = o.rhs(u, 1e-4, ob.LITERAL)
    b = o.rhs(u, 1e-4, ob.FACTORIZED)
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-13
    c = o.step(u, 3, 1e-6, ob.LITERAL)
    d = o.step(u, 3, 1e-6, ob.FACTORIZED)
    assert np.linalg.norm(c - d) / np.linalg.norm(c) < 1e-13


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 3), (3, 1), (3, 2)])
def test_quadratic_face_search_equals_hashed(dim, p):
    m = rect(p) if dim == 2 else box(p)
    a = ob.Oracle(m, 340.0, 1.225, literal_setup=True).connectivity()
    b = ob.Oracle(m, 340.0, 1.225, literal_setup=False).connectivity()
    for k in a:
        assert (a[k] == b[k]).all(), k


def test_face_count_relation():
    for m in (rect(3, 5, 4), box(2, 3)):
        o = ob.Oracle(m, 340.0, 1.225)
        assert 2 * o.F - o.Fb == m.K * o.Nf


def dense_operator(o, n, dt):
    A = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1
        A[:, i] = o.rhs(e.reshape(4, -1), dt, ob.FACTORIZED).ravel()
    return A


@pytest.mark.parametrize("mode", [ob.LITERAL, ob.FACTORIZED])
def test_rk4_is_the_taylor_polynomial(mode):
    m = rect(2, 2, 2)
    o = ob.Oracle(m, 340.0, 1.225, v0=(30.0, 10.0, 0.0))
    n = 4 * m.K * m.Np
    dt = 2e-5
    A = dense_operator(o, n, dt)
    u = np.random.default_rng(3).standard_normal(n)
    exp = u + A @ u + A @ (A @ u) / 2 + A @ (A @ (A @ u)) / 6 + A @ (A @ (A @ (A @ u))) / 24
    got = o.step(u.reshape(4, -1), 1, dt, mode).ravel()
    assert np.linalg.norm(got - exp) / np.linalg.norm(exp) < 1e-13
    eul = o.step(u.reshape(4, -1), 1, dt, mode, euler=True).ravel()
    assert np.linalg.norm(eul - (u + A @ u)) / np.linalg.norm(u) < 1e-13


def mass_and_energy(m, u, c0=340.0, rho=1.225):
    r = dgk.ref_element(m.dim, m.order)
    X = m.node_xyz.reshape(m.K, m.Np, 3)
    if m.dim == 2:
        a = X[:, 1, :2] - X[:, 0, :2]
        b = X[:, 2, :2] - X[:, 0, :2]
        det = np.abs(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0])
    else:
        J = np.stack([X[:, k, :] - X[:, 0, :] for k in (1, 2, 3)], axis=-1)
        det = np.abs(np.linalg.det(J))
    U = u.reshape(4, m.K, m.Np)
    ones = r["M"].sum(axis=0)
    pm = np.sum(det * (U[0] @ ones))
    quad = lambda f: np.einsum("ki,ij,kj->k", f, r["M"], f)
    E = 0.5 * np.sum(det * (quad(U[0]) / (rho * c0 ** 2) + rho * (quad(U[1]) + quad(U[2]) + quad(U[3]))))
    return pm, E


@pytest.mark.parametrize("dim,p", [(2, 3), (3, 2)])
def test_pressure_mass_conserved_energy_nonincreasing(dim, p):
    bcs = "Wall=Reflecting\nAbs=Reflecting\n"
    m = rect(p, 5, 5, bcs) if dim == 2 else box(p, 2, bcs)
    o = ob.Oracle(m, 340.0, 1.225, sign_mode=ob.SIGN_STABLE)
    x = m.node_xyz
    u = np.zeros((4, m.K * m.Np))
    zc = 0.4 if dim == 3 else 0.0
    u[0] = np.exp(-((x[:, 0] - 0.4) ** 2 + (x[:, 1] - 0.45) ** 2 + (x[:, 2] - zc) ** 2) / 0.02)
    u[1] = 0.001 * np.sin(3 * x[:, 1])
    dt = 0.05 * 0.2 / 340.0
    pm0, E0 = mass_and_energy(m, u)
    Eprev = E0
    for _ in range(10):
        u = o.step(u, 5, dt)
        pm, E = mass_and_energy(m, u)
        assert abs(pm - pm0) <= 1e-12 * abs(pm0)
        assert E <= Eprev * (1 + 1e-13)
        Eprev = E
    assert Eprev < E0  # the penalty dissipates


@pytest.mark.parametrize("dim", [2, 3])
def test_constant_state_is_steady_with_reflecting_walls(dim):
    bcs = "Wall=Reflecting\nAbs=Reflecting\n"
    m = rect(3, 3, 3, bcs) if dim == 2 else box(3, 2, bcs)
    o = ob.Oracle(m, 340.0, 1.225)
    u = np.zeros((4, m.K * m.Np))
    u[0] = 2.5
    out = o.rhs(u, 1.0, ob.LITERAL)
    assert np.abs(out).max() < 1e-9 * 2.5 * 340.0
    out = o.rhs(u, 1.0, ob.FACTORIZED)
    assert np.abs(out).max() < 1e-9 * 2.5 * 340.0


def test_reference_sign_rule_per_order():
    """fc*orientation on interior faces: +1 everywhere in 2D, and in 3D at orders where the
    first face point gives an inward normal; recorded here so a change is visible."""
    for p in (1, 2, 3, 4):
        m = box(p)
        o = ob.Oracle(m, 340.0, 1.225)
        c = o.connectivity()
        signs = []
        for f in range(o.F):
            up, dn = c["fNbr"][f]
            if dn < 0:
                continue
            j = list(c["elFIds"][up]).index(f)
            signs.append(o.fc * c["elFOrientation"][up, j])
        signs = set(signs)
        expected = {1: {1}, 2: {1}, 3: {-1}, 4: {1}}[p]
        assert signs == expected, (p, signs)
