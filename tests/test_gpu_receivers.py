"""Receivers: recorded pressure time series against the oracle's state interpolated at the
points with an independent (monomial Vandermonde, physical coordinates) Lagrange basis."""
import numpy as np
import pytest

import dgk
import oracle_binding as ob
from interp import sample
from test_gpu_parity import mixed_mesh, rel

pytestmark = pytest.mark.gpu

EXTENT = {1: (1.0, 0.0, 0.0), 2: (1.0, 0.8, 0.0), 3: (1.0, 0.8, 0.7)}


def points(dim, n, seed):
    rng = np.random.default_rng(seed)
    L = np.array(EXTENT[dim])
    pts = 0.05 * L + 0.9 * L * rng.random((n, 3))
    # near (not on) a face of the structured meshes: DG fields jump across faces, so a point on a
    # face has two valid values and the element choice there is not part of the contract
    pts[0] = 0.5 * L + np.array([1e-3, 7e-4, 3e-4]) * (L > 0)
    return pts


def smooth_state(m, seed=3):
    x = m.node_xyz
    rng = np.random.default_rng(seed)
    c = x.mean(axis=0) + 0.05 * rng.standard_normal(3)
    u = np.zeros((4, m.K * m.Np))
    u[0] = np.exp(-((x - c) ** 2).sum(axis=1) / 0.03)
    u[1] = 0.1 * np.exp(-((x - c) ** 2).sum(axis=1) / 0.05)
    return u


def oracle_series(m, u0, nsteps, dt, pts, v0=(0.0, 0.0, 0.0)):
    o = ob.Oracle(m, 340.0, 1.225, v0=v0, sign_mode=ob.SIGN_STABLE)
    u = u0
    rows = []
    for _ in range(nsteps):
        u = o.step(u, 1, dt)
        rows.append(sample(m, u, pts))
    return np.array(rows)


@pytest.mark.parametrize("dim,p", [(1, 3), (2, 3), (3, 2), (3, 4)])
def test_receivers_match_oracle(dim, p):
    print(f"receivers dim={dim} p={p}", flush=True)
    m = mixed_mesh(dim, p)
    pts = points(dim, 5, 10 + dim)
    u0 = smooth_state(m)
    dt = 2e-6
    s = dgk.Solver(m, 340.0, 1.225, v0=(30.0, 0.0, 0.0), flux_sign=dgk.SIGN_STABLE)
    s.set_state(u0)
    s.set_receivers(pts, 100)
    s.rk4(40, dt)  # one 25-step graph plus 15 direct steps
    rec = s.get_receivers()
    assert rec.shape == (40, len(pts))
    ref = oracle_series(m, u0, 40, dt, pts, v0=(30.0, 0.0, 0.0))
    assert rel(rec, ref) <= 1e-12


def test_receivers_capacity_euler_and_reset():
    m = mixed_mesh(2, 2)
    pts = points(2, 3, 4)
    u0 = smooth_state(m)
    s = dgk.Solver(m, 340.0, 1.225, flux_sign=dgk.SIGN_STABLE)
    s.set_state(u0)
    s.set_receivers(pts, 30)
    s.rk4(40, 1e-6)
    rec = s.get_receivers()
    assert rec.shape == (30, 3)  # rows beyond the capacity are dropped
    s.set_state(u0)
    s.set_receivers(pts, 10)  # resetting clears the record
    s.euler(4, 1e-6)
    e = s.get_receivers()
    assert e.shape == (4, 3)
    o = ob.Oracle(m, 340.0, 1.225, sign_mode=ob.SIGN_STABLE)
    ref = [sample(m, o.step(u0, k + 1, 1e-6, euler=True), pts) for k in range(4)]
    assert rel(e, np.array(ref)) <= 1e-12
    with pytest.raises(dgk.DGKError):
        s.set_receivers([[5.0, 5.0, 0.0]], 10)  # outside the mesh


def test_receivers_fp32_within_bound():
    m = mixed_mesh(3, 3)
    pts = points(3, 4, 6)
    u0 = smooth_state(m)
    out = []
    for prec in (dgk.FP64, dgk.FP32):
        s = dgk.Solver(m, 340.0, 1.225, precision=prec, flux_sign=dgk.SIGN_STABLE)
        s.set_state(u0)
        s.set_receivers(pts, 60)
        s.rk4(60, 2e-6)
        out.append(s.get_receivers())
    assert rel(out[1], out[0]) <= 1e-5


def test_batch_receivers_match_oracle():
    # members 1 and 3 share a 3D group, so member 3's receivers and state sit at an offset
    ms = [mixed_mesh(2, 3), mixed_mesh(3, 2), mixed_mesh(2, 3, 7), mixed_mesh(3, 2, 4)]
    dts = [2e-6, 1.5e-6, 1e-6, 1.2e-6]
    ph = dict(c0=340.0, rho0=1.225, v0=(0.0, 0.0, 0.0))
    b = dgk.Batch(ms, [ph] * 4, dts, flux_sign=dgk.SIGN_STABLE)
    u0 = [smooth_state(m, 5 + i) for i, m in enumerate(ms)]
    for i, u in enumerate(u0):
        b.set_state(i, u)
    pts = {0: points(2, 4, 1), 2: points(2, 2, 2), 1: points(3, 3, 3), 3: points(3, 2, 4)}
    for i, p in pts.items():
        b.set_receivers(i, p, 30 + 5 * i)
    b.rk4(30)
    for i, p in pts.items():
        rec = b.get_receivers(i)
        assert rec.shape == (30, len(p))
        ref = oracle_series(ms[i], u0[i], 30, dts[i], p)
        err = rel(rec, ref)
        print(f"  batch member {i}: receivers rel L2 {err:.2e}", flush=True)
        assert err <= 1e-12


def test_batch_receivers_set_one_member_after_another():
    """Members 1 and 3 share a group record. Member 1's receivers are set, 10 steps run, then
    member 3's are set and 15 more steps run: member 1 keeps all 25 rows, member 3 records the 15
    steps since its receivers were set."""
    ms = [mixed_mesh(2, 3), mixed_mesh(3, 2), mixed_mesh(2, 3, 7), mixed_mesh(3, 2, 4)]
    dts = [2e-6, 1.5e-6, 1e-6, 1.2e-6]
    ph = dict(c0=340.0, rho0=1.225, v0=(0.0, 0.0, 0.0))
    b = dgk.Batch(ms, [ph] * 4, dts, flux_sign=dgk.SIGN_STABLE)
    assert len(b.groups()) == 2
    u0 = [smooth_state(m, 5 + i) for i, m in enumerate(ms)]
    for i, u in enumerate(u0):
        b.set_state(i, u)
    p1, p3 = points(3, 3, 3), points(3, 2, 4)
    b.set_receivers(1, p1, 25)
    b.rk4(10)
    b.set_receivers(3, p3, 15)
    b.rk4(15)
    r1, r3 = b.get_receivers(1), b.get_receivers(3)
    assert r1.shape == (25, 3) and r3.shape == (15, 2)
    ref1 = oracle_series(ms[1], u0[1], 25, dts[1], p1)
    ref3 = oracle_series(ms[3], u0[3], 25, dts[3], p3)[10:]
    assert rel(r1, ref1) <= 1e-12
    assert rel(r3, ref3) <= 1e-12
