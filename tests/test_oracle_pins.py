"""Mathematical pins of the oracle's boundary and mean-flow terms.

The absorbing boundary flux (RKR, src/Mesh.cpp:391-418, 652-667) and the mean-flow part of the
physical flux (src/Mesh.cpp:577-591, entering the Rusanov flux at 519-527 and the reflecting ghost
flux at 616-647) are checked here against things the oracle does not contain:

* an independent assembly of the whole P1 operator in numpy (closed-form P1 mass matrices and
  gradients, vertex averages of the linear flux, face masses) whose absorbing flux is
  1/2 A_n^+ with A_n^+ taken from numpy's eigendecomposition of the v0 = 0 flux Jacobian
  d(n.F)/du -- the reference's RKR is half the characteristic outflow matrix and ignores v0;
* exact solutions with a mean flow: the 1D d'Alembert split of a Gaussian pulse into two pulses
  travelling at v0 +- c, and a Gaussian pressure blob released at rest in 2D (Hankel integral)
  and 3D (closed-form spherical wave), advected by an oblique v0, with observed orders >= p + 1/2;
* invariants: energy does not increase with absorbing or mixed walls (v0 = 0), and a constant
  pressure at rest stays steady with v0 != 0 and reflecting walls;
* the stated time step dt = 0.1 h/c diverges at order 3, while the CFL-0.1 step on the node
  spacing used by the workloads is stable.

Every pin that a planted mistake can reach is also run against mutant builds of oracle.cpp
(absorbing coefficient 1/4 -> 1/2; v0 negated), compiled here with g++ into a temporary
directory, and must reject them (test_mutants_are_rejected).
"""
import os
import subprocess
import types

import numpy as np
import pytest
from scipy.special import j0, j1

import dgk
import oracle_binding as ob
from workloads import conf

C0, RHO = 340.0, 1.225


# ----------------------------------------------------------------------------- independent P1 assembly
def flux_n(n, q, v0):
    """n . F(q) for q = (p, vx, vy, vz) of the linearized Euler equations with mean flow v0."""
    vn = v0 @ n
    p, v = q[0], q[1:]
    return np.concatenate([[vn * p + RHO * C0 * C0 * (v @ n)], vn * v + p / RHO * n])


def half_outflow(n):
    """1/2 A_n^+ of the v0 = 0 Jacobian A_n = d(n.F)/dq, by eigendecomposition."""
    A = np.stack([flux_n(n, e, np.zeros(3)) for e in np.eye(4)], axis=1)
    lam, R = np.linalg.eig(A)
    lam, R = lam.real, R.real
    return 0.5 * R @ np.diag(np.maximum(lam, 0.0)) @ np.linalg.inv(R)


def p1_rhs(mesh, u, dt, v0, face_bc):
    """dt * L(u) on a P1 triangle/tet mesh, assembled from closed forms: M_e = |T| (1 + d_ij) /
    ((d+1)(d+2)); S_i = |T| grad(phi_i) . mean_j F(u_j) (F is linear in u, u is linear); face mass
    |f| (1 + d_ab) / (d (d+1)); interior Rusanov flux with speed c0 (sign +1: P1 under either sign
    rule), reflecting n.F(p, v - (v.n) n), absorbing 1/2 A_n^+ u. face_bc maps the frozenset of a
    boundary face's vertex tags to 1 (reflecting) or 0 (absorbing)."""
    d = mesh.dim
    Np, K = d + 1, mesh.K
    X = mesh.node_xyz.reshape(K, Np, 3)
    T = mesh.node_tags.reshape(K, Np)
    U = u.reshape(4, K, Np)
    v0 = np.asarray(v0, float)
    sides = {}
    for e in range(K):
        for k in range(Np):
            sides.setdefault(frozenset(np.delete(T[e], k).tolist()), []).append(e)
    out = np.zeros_like(U)
    for e in range(K):
        A = np.hstack([np.ones((Np, 1)), X[e, :, :d]])
        grad = np.zeros((Np, 3))
        grad[:, :d] = np.linalg.inv(A)[1:, :].T
        vol = abs(np.linalg.det(A)) / (2 if d == 2 else 6)
        M = vol * (np.ones((Np, Np)) + np.eye(Np)) / ((d + 1) * (d + 2))
        Fm = np.stack([sum(flux_n(nx, U[:, e, j], v0) for j in range(Np)) / Np for nx in np.eye(3)], axis=1)
        S = vol * (Fm @ grad.T)
        B = np.zeros((4, Np))
        for k in range(Np):
            fv = [i for i in range(Np) if i != k]
            a = X[e, fv[0]]
            if d == 2:
                t = X[e, fv[1]] - a
                n = np.array([t[1], -t[0], 0.0])
                area = np.linalg.norm(n)
            else:
                n = np.cross(X[e, fv[1]] - a, X[e, fv[2]] - a)
                area = np.linalg.norm(n) / 2
            n /= np.linalg.norm(n)
            if n @ (a - X[e, k]) < 0:
                n = -n
            Mf = area * (np.ones((d, d)) + np.eye(d)) / (d * (d + 1))
            key = frozenset(np.delete(T[e], k).tolist())
            fl = np.zeros((4, d))
            for ai, i in enumerate(fv):
                qm = U[:, e, i]
                if len(sides[key]) == 2:
                    e2 = [s for s in sides[key] if s != e][0]
                    qp = U[:, e2, list(T[e2]).index(T[e, i])]
                    fl[:, ai] = 0.5 * (flux_n(n, qm, v0) + flux_n(n, qp, v0)) + 0.5 * C0 * (qm - qp)
                elif face_bc[key] == 1:
                    g = qm.copy()
                    g[1:] -= (g[1:] @ n) * n
                    fl[:, ai] = flux_n(n, g, v0)
                else:
                    fl[:, ai] = half_outflow(n) @ qm
            B[:, fv] += fl @ Mf.T
        out[:, e, :] = dt * np.linalg.solve(M, (S - B).T).T
    return out.reshape(4, -1)


def boundary_codes(o, mesh):
    """Boundary face (frozenset of node tags) -> the oracle's BC code (first-node rule)."""
    c = o.connectivity()
    res = {}
    for f in range(o.F):
        up, dn = c["fNbr"][f]
        if dn < 0:
            res[frozenset(mesh.node_tags[up * mesh.Np + c["fNToElNIds"][f, :, 0]].tolist())] = int(c["fBC"][f])
    return res


P1_CASES = [("Wall=Absorbing\nAbs=Absorbing\n", (0.0, 0.0, 0.0)),
            ("Wall=Absorbing\nAbs=Absorbing\n", (40.0, -25.0, 15.0)),
            ("Wall=Reflecting\nAbs=Reflecting\n", (40.0, -25.0, 15.0)),
            ("Wall=Reflecting\nAbs=Absorbing\n", (40.0, -25.0, 15.0))]


def p1_mesh(dim, bcs):
    if dim == 2:
        return dgk.gen_rect(3, 2, 1.0, 0.75, 1, [1, 2, 1, 2], "1:Wall;2:Abs", conf(bcs))
    return dgk.gen_box(2, 2, 2, 1.0, 0.9, 0.8, 1, 0.1, 2, [1, 1, 2, 2, 1, 2], "1:Wall;2:Abs", conf(bcs))


def p1_discrepancy(dim, bcs, v0, library=None):
    m = p1_mesh(dim, bcs)
    v0 = v0 if dim == 3 else (v0[0], v0[1], 0.0)
    o = ob.Oracle(m, C0, RHO, v0=v0, sign_mode=ob.SIGN_STABLE, library=library)
    u = np.random.default_rng(5).standard_normal((4, m.K * m.Np))
    if dim == 2:
        u[3] = 0
    ref = p1_rhs(m, u, 1e-4, v0, boundary_codes(o, m))
    return max(np.linalg.norm(o.rhs(u, 1e-4, mode) - ref) / np.linalg.norm(ref) for mode in (ob.LITERAL, ob.FACTORIZED))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("bcs,v0", P1_CASES)
def test_p1_operator_equals_closed_form_assembly(dim, bcs, v0):
    """Whole operator, both evaluation modes, every boundary kind, with and without mean flow."""
    assert p1_discrepancy(dim, bcs, v0) < 1e-13


def one_element(dim):
    """A single P1 element with no boundary groups: every face is absorbing (src/Mesh.cpp:364-384)."""
    X = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.3, 0.8, 0.0]]) if dim == 2 else \
        np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.05], [0.2, 0.9, 0.1], [0.15, 0.2, 0.7]])
    return types.SimpleNamespace(dim=dim, order=1, K=1, Np=dim + 1, node_xyz=X, node_tags=np.arange(1, dim + 2),
                                 bc_groups=[])


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("v0", [(0.0, 0.0, 0.0), (60.0, -30.0, 20.0)])
def test_absorbing_boundary_term_is_half_the_outflow_characteristic(dim, v0):
    """One element, all faces absorbing: M (rhs/dt) - S(u) = -sum_f E_f M_f (1/2 A_n^+ u) entrywise,
    with S the closed-form P1 volume term; the mean flow does not enter the absorbing flux
    (RKR has no v0, src/Mesh.cpp:397-415)."""
    m = one_element(dim)
    v0 = np.array(v0 if dim == 3 else (v0[0], v0[1], 0.0))
    o = ob.Oracle(m, C0, RHO, v0=tuple(v0))
    rng = np.random.default_rng(11)
    for _ in range(3):
        u = rng.standard_normal((4, m.Np))
        if dim == 2:
            u[3] = 0
        ref = p1_rhs(m, u, 1.0, v0, {frozenset(np.delete(m.node_tags, k).tolist()): 0 for k in range(m.Np)})
        for mode in (ob.LITERAL, ob.FACTORIZED):
            assert np.abs(o.rhs(u, 1.0, mode) - ref).max() <= 1e-12 * np.abs(ref).max()


# ----------------------------------------------------------------------------- exact solutions with mean flow
def dalembert_error(p, n, library=None, v0=50.0, L=2.0, x0=1.0, sig=0.12):
    """1D pulse p0 = exp(-(x-x0)^2/2sig^2), v = 0: p = (p0(x-(v0+c)t) + p0(x-(v0-c)t))/2,
    rho c v = (p0(x-(v0+c)t) - p0(x-(v0-c)t))/2. Energy-norm relative error at a time before the
    pulses reach the (absorbing) ends."""
    T = 0.3 / (C0 + v0)
    m = dgk.gen_line(n, L, p, [1, 1], "1:Abs", conf("Abs=Absorbing\n"))
    o = ob.Oracle(m, C0, RHO, v0=(v0, 0, 0), sign_mode=ob.SIGN_STABLE, library=library)
    x = m.node_xyz[:, 0]
    p0 = lambda s: np.exp(-(s - x0) ** 2 / (2 * sig ** 2))
    u = np.zeros((4, m.K * m.Np))
    u[0] = p0(x)
    h = L / n
    nst = int(np.ceil(T / (0.05 * h / ((C0 + v0) * (p + 1) ** 2))))
    u = o.step(u, nst, T / nst)
    a, b = p0(x - (v0 + C0) * T), p0(x - (v0 - C0) * T)
    M = dgk.ref_element(1, p)["M"] * (h / 2)
    err = nrm = 0.0
    for got, ex in ((u[0], 0.5 * (a + b)), (RHO * C0 * u[1], 0.5 * (a - b))):
        e, f = (got - ex).reshape(m.K, m.Np), ex.reshape(m.K, m.Np)
        err += np.einsum("ki,ij,kj->", e, M, e)
        nrm += np.einsum("ki,ij,kj->", f, M, f)
    return np.sqrt(err / nrm)


D1_MESHES = {1: (40, 80, 160), 2: (20, 40, 80), 3: (10, 20, 40), 4: (10, 20, 40), 5: (8, 16, 32)}


def observed_order(errs):
    return float(np.log2(errs[-2] / errs[-1]))


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_dalembert_split_with_mean_flow_converges(p):
    errs = [dalembert_error(p, n) for n in D1_MESHES[p]]
    assert errs[-1] < errs[-2] < errs[-3]
    assert observed_order(errs) >= p + 0.5, errs


def blob_exact(dim, xr, t, sig):
    """Free-space pressure and radial velocity of p0 = exp(-r^2/2sig^2) released at rest, in the frame
    moving with the mean flow. 3D: p = ((r-ct) g(r-ct) + (r+ct) g(r+ct)) / 2r from the velocity
    potential psi = sig^2 (g(r+ct) - g(r-ct)) / (2 rho c r). 2D: Hankel integrals
    p = int A(k) cos(ckt) J0(kr) k dk, v_r = int A(k) sin(ckt) J1(kr) k dk / (rho c), A = sig^2 e^(-sig^2 k^2/2)."""
    r = np.linalg.norm(xr, axis=1)
    if dim == 3:
        g = lambda s: np.exp(-s ** 2 / (2 * sig ** 2))
        gp = lambda s: -s / sig ** 2 * g(s)
        a, b = r - C0 * t, r + C0 * t
        rs = np.maximum(r, 1e-300)
        p = (a * g(a) + b * g(b)) / (2 * rs)
        vr = sig ** 2 / (2 * RHO * C0) * ((gp(b) - gp(a)) / rs - (g(b) - g(a)) / rs ** 2)
    else:
        k, w = np.polynomial.legendre.leggauss(600)
        kmax = 14.0 / sig
        k, w = 0.5 * kmax * (k + 1), 0.5 * kmax * w
        A = sig ** 2 * np.exp(-(sig * k) ** 2 / 2) * k * w
        kr = np.outer(r, k)
        p = j0(kr) @ (A * np.cos(C0 * k * t))
        vr = j1(kr) @ (A * np.sin(C0 * k * t)) / (RHO * C0)
    return p, vr


def blob_error(dim, p, n, library=None):
    """Gaussian blob in the unit square/cube with an oblique mean flow, stopped well before the wave
    reaches the walls; energy-norm relative error over the whole mesh."""
    if dim == 2:
        v0, ctr, sig, travel = np.array([60.0, -35.0, 0.0]), np.array([0.45, 0.52, 0.0]), 0.06, 0.1
        m = dgk.gen_rect(n, n, 1.0, 1.0, p, [1, 1, 1, 1], "1:Abs", conf("Abs=Absorbing\n"))
    else:
        v0, ctr, sig, travel = np.array([60.0, -35.0, 20.0]), np.array([0.47, 0.52, 0.49]), 0.08, 0.08
        m = dgk.gen_box(n, n, n, 1.0, 1.0, 1.0, p, 0.0, 0, [1] * 6, "1:Abs", conf("Abs=Absorbing\n"))
    T = travel / (C0 + np.linalg.norm(v0))
    x = m.node_xyz
    u = np.zeros((4, m.K * m.Np))
    u[0] = np.exp(-np.sum((x - ctr) ** 2, axis=1) / (2 * sig ** 2))
    o = ob.Oracle(m, C0, RHO, v0=tuple(v0), sign_mode=ob.SIGN_STABLE, library=library, factorized_only=True)
    nst = int(np.ceil(T / (0.1 / n / ((C0 + np.linalg.norm(v0)) * (2 * p + 1)))))
    u = o.step(u, nst, T / nst)
    xr = x - ctr - v0 * T
    pe, vr = blob_exact(dim, xr[:, :2] if dim == 2 else xr, T, sig)
    rr = np.maximum(np.linalg.norm(xr, axis=1), 1e-300)
    ex = [pe] + [vr * xr[:, j] / rr for j in range(3)]
    X = x.reshape(m.K, m.Np, 3)
    J = np.stack([X[:, k, :dim] - X[:, 0, :dim] for k in range(1, dim + 1)], axis=-1)
    det = np.abs(np.linalg.det(J))
    M = dgk.ref_element(dim, p)["M"]
    err = nrm = 0.0
    for q in range(dim + 1):
        w = RHO * C0 if q else 1.0
        e, f = (w * (u[q] - ex[q])).reshape(m.K, m.Np), (w * ex[q]).reshape(m.K, m.Np)
        err += np.sum(det * np.einsum("ki,ij,kj->k", e, M, e))
        nrm += np.sum(det * np.einsum("ki,ij,kj->k", f, M, f))
    return np.sqrt(err / nrm)


BLOB_MESHES = {2: (8, 16, 32), 3: (4, 8, 16)}


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 2), (2, 3), (2, 4), (2, 5), (3, 1), (3, 2), (3, 3)])
def test_gaussian_blob_with_oblique_mean_flow_converges(dim, p):
    errs = [blob_error(dim, p, n) for n in BLOB_MESHES[dim]]
    assert errs[-1] < errs[-2] < errs[-3]
    assert observed_order(errs) >= p + 0.5, errs


# ----------------------------------------------------------------------------- invariants
def energy(m, u):
    """1/2 int (p^2/(rho c^2) + rho |v|^2) with the exact mass matrix of each affine element."""
    X = m.node_xyz.reshape(m.K, m.Np, 3)
    d = m.dim
    J = np.stack([X[:, k, :d] - X[:, 0, :d] for k in range(1, d + 1)], axis=-1)
    det = np.abs(np.linalg.det(J))
    M = dgk.ref_element(d, m.order)["M"]
    U = u.reshape(4, m.K, m.Np)
    quad = lambda f: np.einsum("ki,ij,kj->k", f, M, f)
    return 0.5 * np.sum(det * (quad(U[0]) / (RHO * C0 ** 2) + RHO * (quad(U[1]) + quad(U[2]) + quad(U[3]))))


@pytest.mark.parametrize("dim,p", [(2, 3), (3, 2)])
@pytest.mark.parametrize("bcs", ["Wall=Absorbing\nAbs=Absorbing\n", "Wall=Reflecting\nAbs=Absorbing\n"])
def test_energy_does_not_increase_with_absorbing_walls(dim, p, bcs):
    """v0 = 0: both the Rusanov interior flux and the absorbing flux 1/2 A_n^+ dissipate. The pulse
    sits on the walls from the start, so the boundary terms act throughout."""
    if dim == 2:
        m = dgk.gen_rect(5, 5, 1.0, 0.75, p, [1, 2, 1, 2], "1:Wall;2:Abs", conf(bcs))
    else:
        m = dgk.gen_box(2, 2, 2, 1.0, 0.9, 0.8, p, 0.1, 2, [1, 1, 2, 2, 1, 2], "1:Wall;2:Abs", conf(bcs))
    o = ob.Oracle(m, C0, RHO, sign_mode=ob.SIGN_STABLE)
    x = m.node_xyz
    u = np.zeros((4, m.K * m.Np))
    u[0] = np.exp(-((x[:, 0] - 0.8) ** 2 + (x[:, 1] - 0.1) ** 2 + (x[:, 2] - 0.4 * (dim == 3)) ** 2) / 0.05)
    u[1] = 0.002 * np.cos(2 * x[:, 1])
    dt = 0.05 * 0.2 / C0
    E0 = Eprev = energy(m, u)
    for _ in range(10):
        u = o.step(u, 5, dt)
        E = energy(m, u)
        assert E <= Eprev * (1 + 1e-13)
        Eprev = E
    assert Eprev < 0.99 * E0


@pytest.mark.parametrize("dim", [2, 3])
def test_constant_pressure_at_rest_is_steady_with_mean_flow(dim):
    """u = (p0, 0): the volume term reduces to the boundary integral of n.F(u) = ((v0.n) p0, p0 n/rho),
    which the reflecting ghost flux n.F(p0, 0) cancels exactly; interior jumps vanish."""
    bcs = "Wall=Reflecting\nAbs=Reflecting\n"
    if dim == 2:
        m, v0 = dgk.gen_rect(3, 3, 1.0, 0.75, 3, [1, 2, 1, 2], "1:Wall;2:Abs", conf(bcs)), (60.0, -30.0, 0.0)
    else:
        m, v0 = dgk.gen_box(2, 2, 2, 1.0, 0.9, 0.8, 3, 0.1, 2, [1, 1, 2, 2, 1, 2], "1:Wall;2:Abs", conf(bcs)), \
            (60.0, -30.0, 20.0)
    o = ob.Oracle(m, C0, RHO, v0=v0)
    u = np.zeros((4, m.K * m.Np))
    u[0] = 2.5
    for mode in (ob.LITERAL, ob.FACTORIZED):
        assert np.abs(o.rhs(u, 1.0, mode)).max() < 1e-9 * 2.5 * C0


def test_absorbing_walls_with_mean_flow_are_not_dissipative():
    """Reference behaviour recorded here: the absorbing flux is 1/2 A_n^+ of the v0 = 0 system
    (RKR has no v0), so with a mean flow through the walls it no longer bounds the energy. In the
    unit square at order 3 the energy falls while the pulse is inside, then rises above its start
    once the pulse sits on the walls; with v0 = 0 it only falls."""
    m = dgk.gen_rect(10, 10, 1.0, 1.0, 3, [1, 1, 1, 1], "1:Abs", conf("Abs=Absorbing\n"))
    x = m.node_xyz
    u0 = np.zeros((4, m.K * m.Np))
    u0[0] = np.exp(-((x[:, 0] - 0.5) ** 2 + (x[:, 1] - 0.5) ** 2) / 0.01)
    dt = 0.1 * 0.1 / ((C0 + 50.0) * 7)
    E0 = energy(m, u0)
    for v0, bound in ((50.0, lambda e: max(e) > 2 * E0), (0.0, lambda e: max(e) < E0)):
        o = ob.Oracle(m, C0, RHO, v0=(v0, 0.0, 0.0), sign_mode=ob.SIGN_STABLE)
        u, es = u0, []
        for _ in range(4):
            u = o.step(u, 400, dt)
            es.append(energy(m, u))
        assert bound(es), (v0, [e / E0 for e in es])


def test_stated_time_step_diverges_at_order_3():
    """dt = 0.1 h/c (the CFL 0.1 of the original 2D order-3 case) is beyond the RK4 stability limit of
    the order-3 operator; dt = 0.1 h/(c (2p+1)), the step the workloads use, is stable."""
    n, p = 10, 3
    m = dgk.gen_rect(n, n, 1.0, 1.0, p, [1, 1, 1, 1], "1:Wall", conf("Wall=Reflecting\n"))
    o = ob.Oracle(m, C0, RHO)
    x = m.node_xyz
    u0 = np.zeros((4, m.K * m.Np))
    u0[0] = np.exp(-((x[:, 0] - 0.5) ** 2 + (x[:, 1] - 0.5) ** 2) / 0.005)
    E0 = energy(m, u0)
    h = 1.0 / n
    grown = energy(m, o.step(u0, 50, 0.1 * h / C0))  # 1e17 after 25 steps, 1e43 after 50
    stable = energy(m, o.step(u0, 50 * (2 * p + 1), 0.1 * h / (C0 * (2 * p + 1))))
    assert grown > 1e20 * E0
    assert stable <= E0


# ----------------------------------------------------------------------------- mutants
MUTANTS = {"absorbing_half": "ORC_ABS_COEF=0.5", "v0_negated": "ORC_V0_SIGN=-1.0"}


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")
    src = os.path.join(os.path.dirname(ob.LIB_PATH), "oracle.cpp")
    libs = {}
    for name, define in MUTANTS.items():
        out = str(d / f"liboracle_{name}.so")
        subprocess.run(["g++", "-O2", "-fopenmp", "-fPIC", "-std=c++17", "-shared", "-D" + define, "-o", out, src],
                       check=True)
        libs[name] = ob.load(out)
    return libs


def test_mutants_are_rejected(mutants):
    """Each planted mistake fails the pins that can see it: the P1 closed-form assembly (both), the
    one-element absorbing check (absorbing coefficient), and the exact mean-flow solutions (v0)."""
    for name, L in mutants.items():
        worst = max(p1_discrepancy(dim, bcs, v0, L) for dim in (2, 3) for bcs, v0 in P1_CASES)
        assert worst > 1e-6, name
    L = mutants["absorbing_half"]
    m = one_element(3)
    o = ob.Oracle(m, C0, RHO, library=L)
    u = np.random.default_rng(11).standard_normal((4, m.Np))
    ref = p1_rhs(m, u, 1.0, np.zeros(3), {frozenset(np.delete(m.node_tags, k).tolist()): 0 for k in range(m.Np)})
    assert np.abs(o.rhs(u, 1.0, ob.FACTORIZED) - ref).max() > 1e-3 * np.abs(ref).max()
    L = mutants["v0_negated"]
    errs = [dalembert_error(3, n, L) for n in D1_MESHES[3]]
    assert min(errs) > 0.1 and observed_order(errs) < 1
    errs = [blob_error(2, 2, n, L) for n in BLOB_MESHES[2]]
    assert min(errs) > 0.1 and observed_order(errs) < 1
