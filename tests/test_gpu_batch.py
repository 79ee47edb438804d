"""GPU parity of the batched ensemble: every member against its own CPU-oracle run.

A batch holds simulations of different dimension, order, size, boundary groups, physics
and time step. FP64 members must match the oracle run of the same mesh and dt to
<= 1e-12 relative L2 (the drop-in bar); FP32 members to <= 1e-5.
"""
import numpy as np
import pytest

import dgk
import oracle_binding as ob
from workloads import conf

pytestmark = pytest.mark.gpu

BCS = "Wall=Reflecting\nAbs=Absorbing\n"


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def members():
    """(mesh, physics, dt): uneven sizes and orders, two physics sets, per-member dt."""
    still = dict(c0=340.0, rho0=1.225, v0=(0.0, 0.0, 0.0))
    flow = dict(c0=343.0, rho0=1.2, v0=(30.0, -10.0, 5.0))
    out = [
        (dgk.gen_line(11, 1.0, 3, [1, 2], "1:Wall;2:Abs", conf(BCS)), still, 2e-6),
        (dgk.gen_rect(6, 5, 1.0, 0.8, 2, [1, 2, 1, 2], "1:Wall;2:Abs", conf(BCS)), still, 3e-6),
        (dgk.gen_rect(9, 4, 2.0, 0.7, 2, [2, 2, 1, 1], "1:Wall;2:Abs", conf(BCS)), still, 1.5e-6),
        (dgk.gen_rect(5, 5, 1.0, 1.0, 3, [1, 1, 1, 1], "1:Wall", conf("Wall=Reflecting\n")), flow, 2e-6),
        (dgk.gen_box(3, 2, 2, 1.0, 0.8, 0.7, 2, 0.1, 2, [1, 1, 2, 2, 1, 2], "1:Wall;2:Abs", conf(BCS)), flow, 2.5e-6),
        (dgk.gen_box(2, 3, 2, 0.9, 1.0, 0.6, 2, 0.05, 5, [2, 2, 2, 2, 2, 2], "2:Abs", conf(BCS)), flow, 1e-6),
        (dgk.gen_box(2, 2, 2, 1.0, 1.0, 1.0, 3, 0.1, 3, [1, 2, 1, 2, 1, 2], "1:Wall;2:Abs", conf(BCS)), still, 2e-6),
    ]
    return out


def states(ms, seed=21):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((4, m.K * m.Np)) for m, _, _ in ms]


@pytest.mark.parametrize("schedule", [dgk.SCHEDULE_PARALLEL, dgk.SCHEDULE_SEQUENTIAL])
def test_batch_members_match_oracle_fp64(schedule):
    ms = members()
    b = dgk.Batch([m for m, _, _ in ms], [p for _, p, _ in ms], [d for _, _, d in ms], flux_sign=dgk.SIGN_STABLE)
    g = b.groups()
    assert sum(x["sims"] for x in g) == len(ms)
    # grouped by (dim, order, physics): (1,3) (2,2) (2,3 flow) (3,2 flow) (3,3)
    assert sorted((x["dim"], x["order"], x["sims"]) for x in g) == [(1, 3, 1), (2, 2, 2), (2, 3, 1), (3, 2, 2), (3, 3, 1)]
    u0 = states(ms)
    for i, u in enumerate(u0):
        b.set_state(i, u)
    for i, u in enumerate(u0):  # set/get round trip is exact
        assert np.array_equal(b.get_state(i), u)
    nsteps = 30
    print(f"batch fp64 schedule={schedule} members={len(ms)} steps={nsteps}", flush=True)
    b.rk4(nsteps, schedule)
    for i, (m, ph, dt) in enumerate(ms):
        o = ob.Oracle(m, ph["c0"], ph["rho0"], v0=ph["v0"], sign_mode=ob.SIGN_STABLE)
        ref = o.step(u0[i], nsteps, dt)
        err = rel(b.get_state(i), ref)
        print(f"  member {i} dim={m.dim} p={m.order} K={m.K}: rel L2 {err:.2e}", flush=True)
        assert err <= 1e-12


def test_batch_schedules_are_bitwise_equal_and_chunks_compose():
    ms = members()
    args = ([m for m, _, _ in ms], [p for _, p, _ in ms], [d for _, _, d in ms])
    u0 = states(ms, 5)
    out = []
    for schedule, split in ((dgk.SCHEDULE_PARALLEL, (20,)), (dgk.SCHEDULE_SEQUENTIAL, (7, 13)),
                            (dgk.SCHEDULE_PARALLEL, (13, 7))):
        b = dgk.Batch(*args)
        for i, u in enumerate(u0):
            b.set_state(i, u)
        for n in split:
            b.rk4(n, schedule)
        out.append([b.get_state(i) for i in range(len(ms))])
        b.close()
    for other in out[1:]:
        for a, c in zip(out[0], other):
            assert np.array_equal(a, c)


def test_batch_fp32_members_within_bound():
    ms = members()
    args = ([m for m, _, _ in ms], [p for _, p, _ in ms], [d for _, _, d in ms])
    rng = np.random.default_rng(9)
    u0 = []
    for m, _, _ in ms:  # smooth pulses: the FP32 bound is stated for resolved fields
        x = m.node_xyz
        c = x.mean(axis=0) + 0.1 * rng.standard_normal(3) * (m.dim > 0)
        u = np.zeros((4, m.K * m.Np))
        u[0] = np.exp(-((x - c) ** 2).sum(axis=1) / 0.02)
        u0.append(u)
    b64 = dgk.Batch(*args, precision=dgk.FP64, flux_sign=dgk.SIGN_STABLE)
    b32 = dgk.Batch(*args, precision=dgk.FP32, flux_sign=dgk.SIGN_STABLE)
    for b in (b64, b32):
        for i, u in enumerate(u0):
            b.set_state(i, u)
        b.rk4(60)
    for i in range(len(ms)):
        r = b64.get_state(i)
        g = b32.get_state(i)
        err = rel(g[0], r[0])
        print(f"  member {i}: fp32 p rel L2 {err:.2e}", flush=True)
        assert err <= 1e-5


def test_batch_rejects_bad_input():
    m = dgk.gen_rect(3, 3, 1.0, 1.0, 2, [1, 1, 1, 1], "1:Wall", conf("Wall=Reflecting\n"))
    ph = dict(c0=340.0, rho0=1.225)
    with pytest.raises(dgk.DGKError):
        dgk.Batch([m], [ph], [0.0])
    b = dgk.Batch([m], [ph], [1e-6])
    buf = np.zeros(4 * m.K * m.Np)
    assert dgk.device().dgk_batch_get_state(b.h, 1, dgk._p(buf)) == dgk.DGK_ERR_INVALID  # no member 1
    with pytest.raises(dgk.DGKError):
        b.rk4(-1)


def test_batch_high_order_members():
    """Members at orders 4-5 (3D P5 FP32 runs 8-wave workgroups) in one batch: FP64 against the
    oracle, FP32 against FP64."""
    still = dict(c0=340.0, rho0=1.225, v0=(0.0, 0.0, 0.0))
    ms = [dgk.gen_box(2, 2, 2, 1.0, 1.0, 1.0, 5, 0.1, 3, [1, 2, 1, 2, 1, 2], "1:Wall;2:Abs", conf(BCS)),
          dgk.gen_box(3, 2, 2, 1.0, 0.8, 0.7, 4, 0.1, 4, [1, 1, 2, 2, 1, 2], "1:Wall;2:Abs", conf(BCS)),
          dgk.gen_rect(5, 4, 1.0, 0.8, 5, [1, 2, 1, 2], "1:Wall;2:Abs", conf(BCS)),
          dgk.gen_box(2, 3, 2, 0.9, 1.0, 0.6, 5, 0.05, 5, [2] * 6, "2:Abs", conf(BCS))]
    dts = [1e-6, 1.5e-6, 2e-6, 1.2e-6]
    u0 = []
    for m in ms:
        x = m.node_xyz
        u = np.zeros((4, m.K * m.Np))
        u[0] = np.exp(-((x - x.mean(axis=0)) ** 2).sum(axis=1) / 0.05)
        u0.append(u)
    out = {}
    for prec in (dgk.FP64, dgk.FP32):
        b = dgk.Batch(ms, [still] * len(ms), dts, precision=prec, flux_sign=dgk.SIGN_STABLE)
        for i, u in enumerate(u0):
            b.set_state(i, u)
        b.rk4(40)
        out[prec] = [b.get_state(i) for i in range(len(ms))]
        b.close()
    for i, (m, dt) in enumerate(zip(ms, dts)):
        o = ob.Oracle(m, 340.0, 1.225, sign_mode=ob.SIGN_STABLE)
        ref = o.step(u0[i], 40, dt)
        e64 = rel(out[dgk.FP64][i], ref)
        e32 = rel(out[dgk.FP32][i][0], out[dgk.FP64][i][0])
        print(f"  member {i} dim={m.dim} p={m.order}: fp64 {e64:.2e}, fp32 p {e32:.2e}", flush=True)
        assert e64 <= 1e-12 and e32 <= 1e-5


@pytest.mark.parametrize("precision", [dgk.FP64, dgk.FP32])
def test_batch_group_live_fields_change(precision):
    """Two 2D members in one group: the group's live fields follow its members' states (3 while
    every vz is zero, 4 once one is not), relaying the interleaved group state; each member
    still matches its own oracle run after every change."""
    m0 = dgk.gen_rect(6, 5, 1.0, 0.8, 3, [1, 2, 1, 2], "1:Wall;2:Abs", conf(BCS))
    m1 = dgk.gen_rect(5, 4, 1.0, 0.8, 3, [2, 2, 1, 1], "1:Wall;2:Abs", conf(BCS))
    ph = dict(c0=340.0, rho0=1.225, v0=(20.0, -10.0, 0.0))
    dts = [2e-6, 1.5e-6]
    b = dgk.Batch([m0, m1], [ph, ph], dts, precision=precision, flux_sign=dgk.SIGN_STABLE)
    assert len(b.groups()) == 1
    rng = np.random.default_rng(17)
    u0 = np.zeros((4, m0.K * m0.Np))
    u0[:3] = rng.standard_normal((3, m0.K * m0.Np))        # vz = 0: three live fields
    u1 = rng.standard_normal((4, m1.K * m1.Np))            # vz != 0: four
    u1b = u1.copy()
    u1b[3] = 0.0
    tol = 1e-12 if precision == dgk.FP64 else 1e-5
    oracles = [ob.Oracle(m, ph["c0"], ph["rho0"], v0=ph["v0"], sign_mode=ob.SIGN_STABLE) for m in (m0, m1)]
    b.set_state(0, u0)
    b.set_state(1, u1)  # 3 -> 4 fields: member 0's data is relaid
    b.rk4(10)
    r0 = oracles[0].step(u0, 10, dts[0])
    r1 = oracles[1].step(u1, 10, dts[1])
    assert rel(b.get_state(0), r0) <= tol
    assert rel(b.get_state(1), r1) <= tol
    g0 = b.get_state(0)
    b.set_state(1, u1b)  # 4 -> 3 fields again: member 0's data is relaid back
    assert rel(b.get_state(0), g0) <= (0.0 if precision == dgk.FP64 else 1e-7)
    b.rk4(10)
    assert rel(b.get_state(0), oracles[0].step(r0, 10, dts[0])) <= tol
    got1 = b.get_state(1)
    assert rel(got1, oracles[1].step(u1b, 10, dts[1])) <= tol
    assert not got1[3].any()
    b.close()
