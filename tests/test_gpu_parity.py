"""GPU parity: the HIP path through the C ABI against the CPU oracle.

FP64 bar: relative L2 <= 1e-13 for one operator application and <= 1e-12 after N steps
(the drop-in tolerance). FP32 bar: relative L2 <= 1e-5 against the FP64 oracle.
"""
import numpy as np
import pytest

import dgk
import oracle_binding as ob
import workloads as wl
from workloads import conf

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def mixed_mesh(dim, p, n=None):
    bcs = "Wall=Reflecting\nAbs=Absorbing\n"
    if dim == 1:
        return dgk.gen_line(n or 9, 1.0, p, [1, 2], "1:Wall;2:Abs", conf(bcs))
    if dim == 2:
        return dgk.gen_rect(n or 6, n or 5, 1.0, 0.8, p, [1, 2, 1, 2], "1:Wall;2:Abs", conf(bcs))
    return dgk.gen_box(n or 3, 2, 2, 1.0, 0.8, 0.7, p, 0.1, 2, [1, 1, 2, 2, 1, 2], "1:Wall;2:Abs", conf(bcs))


CASES = [(d, p) for d in (1, 2, 3) for p in (1, 2, 3, 4, 5)]


VARIANTS = [dgk.VARIANT_MFMA, dgk.VARIANT_LDS]


@pytest.mark.parametrize("dim,p", CASES)
@pytest.mark.parametrize("sign", [dgk.SIGN_REFERENCE, dgk.SIGN_STABLE])
@pytest.mark.parametrize("variant", VARIANTS)
def test_rhs_and_rk4_fp64(dim, p, sign, variant):
    print(f"case dim={dim} p={p} sign={sign} variant={variant}", flush=True)
    m = mixed_mesh(dim, p)
    v0 = (40.0, -20.0, 10.0)
    o = ob.Oracle(m, 340.0, 1.225, v0=v0, sign_mode=sign)
    s = dgk.Solver(m, 340.0, 1.225, v0=v0, flux_sign=sign, variant=variant)
    u = np.random.default_rng(11).standard_normal((4, m.K * m.Np))
    assert rel(s.rhs(u, 1e-4), o.rhs(u, 1e-4)) <= 1e-13
    s.set_state(u)
    s.rk4(40, 2e-6)  # one 25-step graph launch plus 15 direct steps
    assert rel(s.get_state(), o.step(u, 40, 2e-6)) <= 1e-12


def test_connectivity_equals_oracle_table():
    for m in (mixed_mesh(2, 3), mixed_mesh(3, 3), wl.cube_3d(4, 2)[0]):
        o = ob.Oracle(m, 340.0, 1.225)
        s = dgk.Solver(m, 340.0, 1.225)
        c = o.connectivity()
        g = s.connectivity()
        r = dgk.ref_element(m.dim, m.order)
        Nf, Nfp = r["fmask"].shape
        for e in range(m.K):
            for j in range(Nf):
                f = c["elFIds"][e, j]
                up, dn = c["fNbr"][f]
                if dn < 0:
                    assert g["nbr"][e, j] == (-2 if c["fBC"][f] == 1 else -1)
                    continue
                other = dn if e == up else up
                assert g["nbr"][e, j] == other
                jn = g["nface"][e, j]
                assert c["elFIds"][other, jn] == f
                # node pairing: oracle maps by tag, device by lattice permutation
                side = 0 if e == up else 1
                pairs_o = {(int(c["fNToElNIds"][f, k, side]), int(c["fNToElNIds"][f, k, 1 - side])) for k in range(Nfp)}
                own = r["fmask"][j]
                tags_e = m.node_tags[e * m.Np:(e + 1) * m.Np]
                tags_n = m.node_tags[other * m.Np:(other + 1) * m.Np]
                pairs_g = set()
                for k in range(Nfp):
                    a = int(own[k])
                    b = int(np.where(tags_n == tags_e[a])[0][0])
                    pairs_g.add((a, b))
                assert pairs_o == pairs_g
                # interior penalty sign: fc * orientation of the lower-index element
                jup = list(c["elFIds"][up]).index(f)
                assert g["sign"][e, j] == o.fc * c["elFOrientation"][up, jup]


def test_square_2d_1000_steps_fp64():
    """The 2D P3 workload (3200 elements, reflecting walls, Gaussian pulse) for 1000 RK4 steps."""
    m, ph, dt, u0 = wl.square_2d()
    o = ob.Oracle(m, ph["c0"], ph["rho0"], factorized_only=True)
    s = dgk.Solver(m, ph["c0"], ph["rho0"])
    s.set_state(u0)
    s.rk4(1000, dt)
    g = s.get_state()
    c = o.step(u0, 1000, dt)
    for q in range(3):
        assert rel(g[q], c[q]) <= 1e-12, q
    assert np.abs(g[3]).max() == 0.0


def test_cube_3d_small_both_signs_fp64():
    for sign in (dgk.SIGN_REFERENCE, dgk.SIGN_STABLE):
        m, ph, dt, u0 = wl.cube_3d(6, 3)
        o = ob.Oracle(m, ph["c0"], ph["rho0"], v0=ph["v0"], sign_mode=sign)
        s = dgk.Solver(m, ph["c0"], ph["rho0"], v0=ph["v0"], flux_sign=sign)
        s.set_state(u0)
        s.rk4(200, dt)
        assert rel(s.get_state(), o.step(u0, 200, dt)) <= 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_euler_and_sources_fp64(dim):
    m = mixed_mesh(dim, 2, 8 if dim == 2 else 4)
    o = ob.Oracle(m, 340.0, 1.225, v0=(10.0, 0.0, 0.0))
    s = dgk.Solver(m, 340.0, 1.225, v0=(10.0, 0.0, 0.0))
    z = 0.3 if dim == 3 else 0.0
    src = np.array([[0, 0.5, 0.4, z, 0.2, 1.0, 800.0, 0.3, 1e-4],
                    [1, 0.3, 0.4, z, 0.1, 2.0, 500.0, 0.0, 1.0],
                    [1, 0.5, 0.4, z, 0.1, 2.0, 500.0, np.pi, 1.0]])
    dt = 1e-6
    ts = np.cumsum(np.full(60, dt)) - dt
    u0 = np.zeros((4, m.K * m.Np))
    s.set_sources(src)
    s.set_state(u0)
    s.rk4(60, dt, ts)
    assert rel(s.get_state(), o.step(u0, 60, dt, t_steps=ts, sources=src)) <= 1e-12
    s.set_state(u0)
    s.euler(30, dt, ts[:30])
    assert rel(s.get_state(), o.step(u0, 30, dt, t_steps=ts[:30], sources=src, euler=True)) <= 1e-12


@pytest.mark.parametrize("prec", [dgk.FP64, dgk.FP32])
def test_sources_replayed_from_graphs_across_calls(prec):
    """Source steps run from captured graphs (the source kernels read the step's start time on the
    device): calls of 7, 25, 1 and 40 steps in a row, a source that ends inside the second call and
    one that lasts, against one oracle run over all 73 steps."""
    print(f"sources through graphs, precision {prec}", flush=True)
    m = mixed_mesh(2, 3, 8)
    o = ob.Oracle(m, 340.0, 1.225, v0=(5.0, 0.0, 0.0))
    s = dgk.Solver(m, 340.0, 1.225, v0=(5.0, 0.0, 0.0), precision=prec)
    dt = 1e-6
    ts = np.cumsum(np.full(73, dt)) - dt
    src = np.array([[0, 0.5, 0.4, 0.0, 0.2, 1.0, 800.0, 0.3, ts[20] + 0.5 * dt],
                    [0, 0.3, 0.3, 0.0, 0.15, 2.0, 500.0, 0.0, 1.0]])
    u0 = np.zeros((4, m.K * m.Np))
    s.set_sources(src)
    s.set_state(u0)
    i = 0
    for n in (7, 25, 1, 40):
        s.rk4(n, dt, ts[i:i + n])
        i += n
    ref = o.step(u0, 73, dt, t_steps=ts, sources=src)
    assert rel(s.get_state(), ref) <= (1e-12 if prec == dgk.FP64 else 1e-5)


def test_cube_3d_full_size_fp64():
    """The 3D P3 workload at full size (196,608 tets): 3 steps against the oracle, then finite."""
    m, ph, dt, u0 = wl.cube_3d()
    assert m.K == 196608
    s = dgk.Solver(m, ph["c0"], ph["rho0"], v0=ph["v0"], flux_sign=dgk.SIGN_STABLE)
    o = ob.Oracle(m, ph["c0"], ph["rho0"], v0=ph["v0"], sign_mode=ob.SIGN_STABLE, factorized_only=True)
    s.set_state(u0)
    s.rk4(3, dt)
    assert rel(s.get_state(), o.step(u0, 3, dt)) <= 1e-13
    assert s.check_finite()


def test_nonfinite_state_is_reported():
    m = mixed_mesh(2, 1)
    s = dgk.Solver(m, 340.0, 1.225)
    u = np.zeros((4, m.K * m.Np))
    s.set_state(u)
    assert s.check_finite()
    u[1, 5] = np.nan
    s.set_state(u)
    assert not s.check_finite()


def test_nonconforming_mesh_is_rejected():
    m = mixed_mesh(2, 2)
    tags = m.node_tags.copy()
    # element 0 is the lower triangle of the first cell; its local edge (1,2) is interior. Give that
    # edge's node (local node 4 at order 2) a tag nobody shares: the face no longer matches its neighbour
    tags[4] = tags.max() + 1
    m.node_tags = tags
    with pytest.raises(dgk.DGKError) as ei:
        dgk.Solver(m, 340.0, 1.225)
    assert ei.value.status == 3


@pytest.mark.parametrize("dim,p", [(2, 3), (3, 2), (3, 3), (3, 4), (3, 5)])
@pytest.mark.parametrize("variant", VARIANTS)
def test_rhs_fp32(dim, p, variant):
    m = mixed_mesh(dim, p)
    v0 = (40.0, -20.0, 10.0)
    o = ob.Oracle(m, 340.0, 1.225, v0=v0, sign_mode=ob.SIGN_STABLE)
    s = dgk.Solver(m, 340.0, 1.225, v0=v0, flux_sign=dgk.SIGN_STABLE, precision=dgk.FP32, variant=variant)
    u = np.random.default_rng(5).standard_normal((4, m.K * m.Np))
    assert rel(s.rhs(u, 1e-4), o.rhs(u, 1e-4)) <= 2e-6


def _fp32_run(workload, fp32_flags):
    if workload == "square":
        m, ph, dt, u0 = wl.square_2d()
        n = 1000
    else:
        m, ph, dt, u0 = wl.cube_3d(8, 3)
        n = 300
    o = ob.Oracle(m, ph["c0"], ph["rho0"], v0=ph["v0"], sign_mode=ob.SIGN_STABLE)
    s = dgk.Solver(m, ph["c0"], ph["rho0"], v0=ph["v0"], precision=dgk.FP32, flux_sign=dgk.SIGN_STABLE,
                   fp32_flags=fp32_flags)
    s.set_state(u0)
    s.rk4(n, dt)
    g = s.get_state()
    c = o.step(u0, n, dt)
    return [rel(g[q], c[q]) for q in range(m.dim + 1)]


FP32_FORMS = {"compensated": 0, "plain": dgk.FP32_PLAIN_STATE, "corrected": dgk.FP32_OPERATOR_CORRECTION}


@pytest.mark.parametrize("workload", ["square", "cube"])
@pytest.mark.parametrize("form", list(FP32_FORMS))
def test_fp32_error_bound(workload, form):
    errs = _fp32_run(workload, FP32_FORMS[form])
    print("fp32 rel-L2 per field", workload, form, errs)
    assert max(errs) <= 1e-5


def test_fp32_compensated_state_removes_accumulated_rounding():
    plain = max(_fp32_run("square", dgk.FP32_PLAIN_STATE))
    comp = max(_fp32_run("square", 0))
    print(f"square 1000 steps: plain {plain:.3e} compensated {comp:.3e}")
    assert comp < 0.5 * plain


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_fp64_state_round_trip_is_exact(dim):
    """set_state/get_state move FP64 values unchanged, through field planes (1D/2D) or the
    node-interleaved 3D device layout, also for batch members at a group offset."""
    m = mixed_mesh(dim, 3)
    u = np.random.default_rng(9).standard_normal((4, m.K * m.Np))
    if dim < 3:
        u[dim + 1:] = 0.0
    s = dgk.Solver(m, 340.0, 1.225)
    s.set_state(u)
    assert np.array_equal(s.get_state(), u)
    s.close()
    m2 = mixed_mesh(dim, 3, 5)
    u2 = np.random.default_rng(10).standard_normal((4, m2.K * m2.Np))
    ph = dict(c0=340.0, rho0=1.225, v0=(0.0, 0.0, 0.0))
    b = dgk.Batch([m, m2], [ph, ph], [1e-6, 2e-6])
    b.set_state(1, u2)
    b.set_state(0, u)
    assert np.array_equal(b.get_state(0), u)
    assert np.array_equal(b.get_state(1), u2)
    b.close()


def test_fp32_compensated_state_round_trip():
    m = mixed_mesh(3, 3)
    s = dgk.Solver(m, 340.0, 1.225, precision=dgk.FP32)
    u = np.random.default_rng(8).standard_normal((4, m.K * m.Np))
    s.set_state(u)
    back = s.get_state()
    assert rel(back, u) <= 2.0 ** -30  # FP32 value plus bf16 low word
    p = dgk.Solver(m, 340.0, 1.225, precision=dgk.FP32, fp32_flags=dgk.FP32_PLAIN_STATE)
    p.set_state(u)
    assert 2.0 ** -27 <= rel(p.get_state(), u) <= 2.0 ** -23


@pytest.mark.parametrize("dim", [2, 3])
def test_fp32_euler_and_sources_compensated(dim):
    m = mixed_mesh(dim, 3, 8 if dim == 2 else 4)
    o = ob.Oracle(m, 340.0, 1.225, sign_mode=ob.SIGN_STABLE)
    s = dgk.Solver(m, 340.0, 1.225, precision=dgk.FP32, flux_sign=dgk.SIGN_STABLE)
    z = 0.3 if dim == 3 else 0.0
    u0 = wl.gaussian(m, (0.5, 0.4, z), 0.01)
    src = np.array([[0, 0.3, 0.3, z, 0.1, 0.5, 2000.0, 0.2, 4e-5]])
    dt = 1e-6
    ts = np.arange(60) * dt
    s.set_sources(src)
    s.set_state(u0)
    s.rk4(60, dt, ts)
    s.euler(10, dt, ts[:10] + 60 * dt)
    s.rk4(30, dt, ts[:30] + 70 * dt)
    ref = o.step(u0, 60, dt, t_steps=ts, sources=src)
    ref = o.step(ref, 10, dt, t_steps=ts[:10] + 60 * dt, sources=src, euler=True)
    ref = o.step(ref, 30, dt, t_steps=ts[:30] + 70 * dt, sources=src)
    assert rel(s.get_state()[:dim + 1], ref[:dim + 1]) <= 1e-5


@pytest.mark.parametrize("dim", [1, 2])
def test_zero_velocity_planes_stay_zero_and_switch(dim):
    """1D/2D states whose extra velocity components are zero run on dim+1 planes; a later state
    with nonzero extra components (and back) gives the oracle's result either way."""
    m = mixed_mesh(dim, 3)
    o = ob.Oracle(m, 340.0, 1.225, v0=(20.0, 0.0, 0.0), sign_mode=ob.SIGN_STABLE)
    s = dgk.Solver(m, 340.0, 1.225, v0=(20.0, 0.0, 0.0), flux_sign=dgk.SIGN_STABLE)
    rng = np.random.default_rng(3)
    full = rng.standard_normal((4, m.K * m.Np))
    live = full.copy()
    live[dim + 1:] = 0.0
    for u in (live, full, live):
        s.set_state(u)
        s.euler(3, 2e-6)
        s.rk4(30, 2e-6)
        ref = o.step(o.step(u, 3, 2e-6, euler=True), 30, 2e-6)
        g = s.get_state()
        assert rel(g, ref) <= 1e-12
        if u is live:
            assert not g[dim + 1:].any()
        assert rel(s.rhs(u, 1e-4), o.rhs(u, 1e-4)) <= 1e-13


def test_room_full_size_fp32_vs_fp64_with_receivers():
    """The 20 x 15 x 8 m room with its block obstacle at order 4 (920,064 tets, 129 M unknowns):
    the FP32 form against the FP64 drop-in after 20 steps, fields and 8 receivers."""
    m, ph, dt, u0 = wl.room_3d()
    assert m.K == 920064 and m.Np == 35
    rng = np.random.default_rng(6)
    center = m.node_xyz[np.argmax(u0[0])]  # receivers around the pulse, which moves ~0.02 m in 20 steps
    rec = center + 0.15 * (rng.random((8, 3)) - 0.5)
    out = {}
    for prec in (dgk.FP64, dgk.FP32):
        s = dgk.Solver(m, ph["c0"], ph["rho0"], precision=prec, flux_sign=dgk.SIGN_STABLE)
        s.set_state(u0)
        s.set_receivers(rec, 20)
        s.rk4(20, dt)
        out[prec] = (s.get_state(), s.get_receivers())
        assert s.check_finite()
        s.close()
    g64, g32 = out[dgk.FP64][0], out[dgk.FP32][0]
    assert rel(g32[0], g64[0]) <= 1e-5
    assert rel(out[dgk.FP32][1], out[dgk.FP64][1]) <= 1e-4


@pytest.mark.parametrize("dim,p", CASES)
def test_fp32_rk4_every_shape_against_oracle(dim, p):
    """FP32 form on every (dim, order): the interleaved state with 2, 3 and 4 fields, operators
    whose rows all sit in the 4x4 tail MFMA (1D orders 1-3, 2D/3D order 1), tails beside 16x16x4
    blocks (3D orders 3 and 4) and the 8-wave workgroups of 3D order 5."""
    m = mixed_mesh(dim, p, 4 if dim == 3 else None)
    x = m.node_xyz
    u0 = np.zeros((4, m.K * m.Np))
    u0[0] = np.exp(-((x - x.mean(axis=0)) ** 2).sum(axis=1) / 0.05)
    o = ob.Oracle(m, 340.0, 1.225, v0=(20.0, 10.0, 0.0), sign_mode=ob.SIGN_STABLE)
    s = dgk.Solver(m, 340.0, 1.225, v0=(20.0, 10.0, 0.0), precision=dgk.FP32, flux_sign=dgk.SIGN_STABLE)
    s.set_state(u0)
    s.rk4(60, 1e-6)
    ref = o.step(u0, 60, 1e-6)
    g = s.get_state()
    errs = [rel(g[q], ref[q]) for q in range(dim + 1)]
    print(f"fp32 dim={dim} p={p}: {errs}")
    assert max(errs) <= 1e-5
