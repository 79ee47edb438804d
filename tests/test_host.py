"""Host side of the boundary: config semantics, MSH ingestion, generators, BC rule."""
import os
import struct
import subprocess

import numpy as np
import pytest

import dgk
import oracle_binding as ob
from workloads import conf


# ----------------------------------------------------------------------------- config
def test_config_semantics():
    text = ("[Config]\n# comment line\n  timeStart = 0\ntimeEnd=0.1\ntimeStep=1e-4\ntimeRate=0.001\n"
            "elementType = Lagrange\ntimeIntMethod=Runge-Kutta\nnumThreads=1\nv0_x = 1\nv0_y=2\nv0_z=3\n"
            "rho0=1.225\nc0 = 343.3\nsaveFile = out file.msh\n"
            "source1 = monopole, 0,0,0, 0.5, 1,100,0,0.05\n"
            "source2 = dipole, 1,2,3, 0.2, 2,50,0.5,1\n"
            "source3 = quadrupole, 0,0,0, 0.4, 3,10,0,1\n"
            "initialCondtition1 = gaussian, 0.5,0.5,0,0.005,1\n"
            "timeEnd=0.2\n")
    c = dgk.parse_config(text)
    assert c["timeEnd"] == 0.2  # later duplicate wins (std::map assignment)
    assert c["numThreads"] == 0  # 1 maps to 0 (src/configParser.cpp:58)
    assert c["saveFile"] == "outfile.msh"  # all whitespace removed
    assert c["v0"] == [1, 2, 3] and c["c0"] == 343.3
    s = c["sources"]
    assert s[0] == [0, 0, 0, 0, 0.5, 1, 100, 0, 0.05]
    assert s[1] == [1, 1 - 0.2, 2, 3, 0.1, 2, 50, 0.5, 1]
    assert s[2] == pytest.approx([1, 1 + 0.2, 2, 3, 0.1, 2, 50, 0.5 + np.pi, 1])
    assert [q[1:4] for q in s[3:7]] == [[-0.4, 0, 0], [0.4, 0, 0], [0, -0.4, 0], [0, 0.4, 0]]
    assert [q[7] for q in s[3:7]] == pytest.approx([0, 0, np.pi, np.pi])
    assert c["initConditions"] == [[0, 0.5, 0.5, 0, 0.005, 1]]


def test_config_missing_required_key_is_an_error():
    text = "timeStart=0\ntimeEnd=1\ntimeStep=0.1\ntimeRate=1\nnumThreads=2\nv0_x=0\nv0_y=0\nv0_z=0\nrho0=1\n"
    with pytest.raises(ValueError, match="c0"):
        dgk.parse_config(text)
    with pytest.raises(ValueError, match="gpu.precision"):
        dgk.parse_config(text + "c0=1\ngpu.precision=fp16\n")


def test_reference_config_files(reference_dir):
    c = dgk.parse_config(open(os.path.join(reference_dir, "doc/config/config.conf")).read())
    assert (c["timeStart"], c["timeEnd"], c["timeStep"], c["timeRate"]) == (0, 0.1, 0.0001, 0.001)
    assert c["numThreads"] == 4 and c["c0"] == 343.3 and c["rho0"] == 1.225
    assert c["sources"] == [] and c["initConditions"] == [[0, 0, 0, 0, 1, 1]]
    c = dgk.parse_config(open(os.path.join(reference_dir, "doc/config/Room_2D.conf")).read())
    assert c["sources"] == [[0, 3, 1.2, 0, 0.1, 10, 1500, 0, 0.01]]
    c = dgk.parse_config(open(os.path.join(reference_dir, "doc/config/Amphi_Pulse_3D.conf")).read())
    assert c["initConditions"] == [[0, 6.5, 6.5, 1.8, 1, 0.05]] and c["timeStep"] == 1e-5


# ----------------------------------------------------------------------------- MSH
@pytest.mark.parametrize("rel,dim,K", [("doc/1d/line.msh", 1, 100), ("doc/2d/square.msh", 2, None),
                                       ("doc/3d/cube.msh", 3, None), ("doc/2d/room.msh", 2, None)])
def test_read_reference_meshes_and_roundtrip(reference_dir, tmp_path, rel, dim, K):
    cfg = conf("Absorbing=Absorbing\nReflecting=Reflecting\nBoundary=Absorbing\n")
    m = dgk.read_mesh(os.path.join(reference_dir, rel), cfg)
    assert m.dim == dim and m.order == 1
    if K:
        assert m.K == K
    X = m.node_xyz.reshape(m.K, m.Np, 3)
    assert np.isfinite(X).all()
    out = str(tmp_path / "rt.msh")
    m.write(out)
    m2 = dgk.read_mesh(out, cfg)
    assert (m2.node_tags == m.node_tags).all() and np.array_equal(m2.node_xyz, m.node_xyz)
    assert [(g[0], g[1], len(g[2])) for g in m2.bc_groups] == [(g[0], g[1], len(g[2])) for g in m.bc_groups]
    # the oracle builds a consistent face table on the reference's own mesh
    o = ob.Oracle(m, 343.0, 1.225)
    assert 2 * o.F - o.Fb == m.K * o.Nf


def _two_triangles_v22(path, binary=False):
    nodes = [(1, 0.0, 0.0), (2, 1.0, 0.0), (3, 1.0, 1.0), (4, 0.0, 1.0)]
    with open(path, "w") as f:
        f.write('$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$PhysicalNames\n2\n1 7 "Wall"\n2 9 "Domain"\n'
                "$EndPhysicalNames\n$Nodes\n4\n")
        for t, x, y in nodes:
            f.write(f"{t} {x} {y} 0\n")
        f.write("$EndNodes\n$Elements\n6\n")
        f.write("1 1 2 7 1 1 2\n2 1 2 7 1 2 3\n3 1 2 7 1 3 4\n4 1 2 7 1 4 1\n")
        f.write("5 2 2 9 5 1 2 3\n6 2 2 9 5 1 3 4\n$EndElements\n")


def _two_triangles_v41_binary(path):
    b = bytearray()
    b += b"$MeshFormat\n4.1 1 8\n" + struct.pack("<i", 1) + b"\n$EndMeshFormat\n"
    b += b'$PhysicalNames\n2\n1 7 "Wall"\n2 9 "Domain"\n$EndPhysicalNames\n'
    b += b"$Entities\n" + struct.pack("<4Q", 0, 1, 1, 0)
    b += struct.pack("<i6dQiQ", 1, 0, 0, 0, 1, 1, 0, 1, 7, 0)
    b += struct.pack("<i6dQiQ", 5, 0, 0, 0, 1, 1, 0, 1, 9, 0)
    b += b"\n$EndEntities\n$Nodes\n" + struct.pack("<4Q", 1, 4, 1, 4) + struct.pack("<iiiQ", 2, 5, 0, 4)
    b += struct.pack("<4Q", 1, 2, 3, 4) + struct.pack("<12d", 0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0)
    b += b"\n$EndNodes\n$Elements\n" + struct.pack("<4Q", 2, 6, 1, 6)
    b += struct.pack("<iiiQ", 1, 1, 1, 4) + struct.pack("<12Q", 1, 1, 2, 2, 2, 3, 3, 3, 4, 4, 4, 1)
    b += struct.pack("<iiiQ", 2, 5, 2, 2) + struct.pack("<8Q", 5, 1, 2, 3, 6, 1, 3, 4)
    b += b"\n$EndElements\n"
    open(path, "wb").write(bytes(b))


@pytest.mark.parametrize("fmt", ["v22", "v41b"])
def test_other_msh_formats(tmp_path, fmt):
    p = str(tmp_path / "m.msh")
    _two_triangles_v22(p) if fmt == "v22" else _two_triangles_v41_binary(p)
    m = dgk.read_mesh(p, conf("Wall=Reflecting\n"))
    assert (m.dim, m.order, m.K) == (2, 1, 2)
    assert m.node_tags.tolist() == [1, 2, 3, 1, 3, 4]
    assert m.bc_groups[0][0] == 7 and m.bc_groups[0][1] == 1 and m.bc_groups[0][2].tolist() == [1, 2, 3, 4]
    assert (m.node_bc == 1).all()


def test_truncated_mesh_is_an_error(tmp_path, reference_dir):
    src = open(os.path.join(reference_dir, "doc/2d/square.msh")).read()
    p = tmp_path / "t.msh"
    p.write_text(src[: len(src) // 2])
    with pytest.raises(RuntimeError):
        dgk.read_mesh(str(p), conf())


# ----------------------------------------------------------------------------- generators
def test_generator_sizes_and_orientation():
    m = dgk.gen_rect(40, 40, 1.0, 1.0, 3, [1, 1, 1, 1], "1:Wall", conf("Wall=Reflecting\n"))
    assert (m.K, m.Np) == (3200, 10)
    X = m.node_xyz.reshape(m.K, m.Np, 3)
    a, b = X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]
    assert (a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0] > 0).all()
    m = dgk.gen_box(4, 3, 2, 1.0, 1.0, 1.0, 3, 0.1, 2, [1] * 6, "1:Abs", conf("Abs=Absorbing\n"))
    assert (m.K, m.Np) == (6 * 24, 20)
    X = m.node_xyz.reshape(m.K, m.Np, 3)
    J = np.stack([X[:, k] - X[:, 0] for k in (1, 2, 3)], axis=-1)
    assert (np.linalg.det(J) > 0).all()
    # shared high-order nodes carry one tag and one position
    tags, idx = np.unique(m.node_tags, return_index=True)
    pos = {t: m.node_xyz[i] for t, i in zip(tags, idx)}
    for t, x in zip(m.node_tags, m.node_xyz):
        assert np.array_equal(pos[t], x)


def test_room_block_count():
    block = ((4, 4, 0), (12, 12, 4))
    m = dgk.gen_box(20, 15, 8, 5.0, 3.75, 2.0, 1, 0.0, 0, [1, 1, 1, 1, 1, 2], "1:Walls;2:Ceiling",
                    conf("Walls=Reflecting\nCeiling=Absorbing\n"), block=block, block_tag=1)
    assert m.K == 6 * (20 * 15 * 8 - 8 * 8 * 4)


def test_first_node_bc_rule_matches_oracle():
    """Corner nodes belong to both groups: the later (higher-tag) group wins, and a face's BC
    is read at its first node (src/Mesh.cpp:364-384)."""
    for order in (1, 2):
        m = dgk.gen_rect(3, 3, 1.0, 1.0, order, [1, 2, 1, 2], "1:Wall;2:Abs",
                         conf("Wall=Reflecting\nAbs=Absorbing\n"))
        o = ob.Oracle(m, 340.0, 1.225)
        c = o.connectivity()
        r = dgk.ref_element(2, order)
        got = []
        for e in range(m.K):
            for j in range(3):
                f = c["elFIds"][e, j]
                if c["fBC"][f] >= 0:
                    got.append((int(c["fBC"][f]), int(m.node_bc[e * m.Np + r["fmask"][j, 0]])))
        assert got and all(a == b for a, b in got)
        assert {a for a, _ in got} == {0, 1}


def test_first_node_bc_rule_by_hand():
    """Boundary codes of a 2 x 1 rectangle derived from coordinates alone: the sides y = 0 and
    y = 1 are in group 1 (Wall, reflecting), x = 0 and x = 1 in group 2 (Abs, absorbing). A
    boundary face takes the code of its first node; a corner node lies in both groups, and the
    group applied last (ascending physical tag: Abs) wins (src/Mesh.cpp:364-384). So a face on a
    Wall side is reflecting exactly when its first node is not a corner, and every face on an Abs
    side is absorbing. Checked against the oracle's face table and the host's per-node codes."""
    for order in (1, 2, 3):
        m = dgk.gen_rect(2, 1, 1.0, 1.0, order, [1, 2, 1, 2], "1:Wall;2:Abs",
                         conf("Wall=Reflecting\nAbs=Absorbing\n"))
        c = ob.Oracle(m, 340.0, 1.225).connectivity()
        r = dgk.ref_element(2, order)
        xyz = m.node_xyz.reshape(m.K, m.Np, 3)
        on = lambda v, a: abs(v - a) < 1e-12  # noqa: E731
        seen = {0: 0, 1: 0}
        wall_corner = 0
        for e in range(m.K):
            for j in range(3):
                a, b = xyz[e, r["fmask"][j, 0]], xyz[e, r["fmask"][j, -1]]
                wall_side = (on(a[1], 0) and on(b[1], 0)) or (on(a[1], 1) and on(b[1], 1))
                abs_side = (on(a[0], 0) and on(b[0], 0)) or (on(a[0], 1) and on(b[0], 1))
                f = c["elFIds"][e, j]
                if not (wall_side or abs_side):
                    assert c["fBC"][f] < 0  # interior face
                    continue
                corner = (on(a[0], 0) or on(a[0], 1)) and (on(a[1], 0) or on(a[1], 1))
                want = 1 if wall_side and not corner else 0
                wall_corner += wall_side and corner
                assert c["fBC"][f] == want, (order, e, j, a, b)
                assert m.node_bc[e * m.Np + r["fmask"][j, 0]] == want
                seen[want] += 1
        assert seen[0] > 0 and seen[1] > 0  # both outcomes occur on this mesh
        assert wall_corner > 0  # and the corner rule decides some Wall-side face


# ----------------------------------------------------------------------------- CLI
def test_cli_wrong_argc_returns_e2big():
    exe = os.path.join(dgk.PKG_DIR, "dgalerkin")
    if not os.path.exists(exe):
        pytest.skip("dgalerkin not built")
    r = subprocess.run([exe], capture_output=True)
    assert r.returncode == 7  # E2BIG


@pytest.mark.parametrize("dim,p", [(1, 3), (2, 2), (2, 5), (3, 3), (3, 5)])
def test_interpolation_helper_reproduces_polynomials(dim, p):
    """The receiver tests' independent interpolation is exact for degree-p polynomials."""
    from interp import sample
    from workloads import conf

    bcs = conf("Wall=Reflecting\n")
    if dim == 1:
        m = dgk.gen_line(5, 1.0, p, [1, 1], "1:Wall", bcs)
    elif dim == 2:
        m = dgk.gen_rect(3, 2, 1.0, 0.8, p, [1, 1, 1, 1], "1:Wall", bcs)
    else:
        m = dgk.gen_box(2, 2, 1, 1.0, 0.8, 0.7, p, 0.1, 3, [1] * 6, "1:Wall", bcs)
    rng = np.random.default_rng(1)
    coef = rng.standard_normal(4)

    def f(x):
        s = coef[0] + coef[1] * x[..., 0]
        if dim > 1:
            s = s + coef[2] * x[..., 1] ** 2 * x[..., 0] ** (p - 2 if p > 2 else 0)
        if dim > 2:
            s = s + coef[3] * x[..., 2] ** p
        return s + x[..., 0] ** p

    u = np.zeros((4, m.K * m.Np))
    u[0] = f(m.node_xyz)
    ext = np.array([1.0, 0.8 if dim > 1 else 0.0, 0.7 if dim > 2 else 0.0])
    pts = 0.05 * ext + 0.9 * ext * rng.random((6, 3))
    assert np.allclose(sample(m, u, pts), f(pts), rtol=1e-10, atol=1e-10)


def test_batch_group_tuning_choice_rule():
    """Batch.tune_groups' host logic, with the device calls stubbed: every candidate is set on the
    groups that have a kernel of that shape (others refuse with DGK_ERR_UNSUPPORTED and are not
    timed on it), each group keeps its fastest candidate, one of another tile shape than the
    built-in one only when it is below 0.98 of the built-in shape's best, and the choice is left
    set on the group."""
    # stage us per (group, shape, grid, order); group 1 has no kernel of shape 3
    table = {0: {0: 100.0, 1: 97.0, 3: 99.5}, 1: {0: 50.0, 1: 49.5, 2: 40.0}}
    grid_gain = {100: 0.0, 200: -1.0}
    b = object.__new__(dgk.Batch)
    state = {0: [0, 100, 0], 1: [0, 100, 0]}
    calls = []

    def set_tile_shape(gi, sh):
        if sh not in table[gi]:
            raise dgk.DGKError(dgk.DGK_ERR_UNSUPPORTED, "no kernel")
        state[gi][0] = sh

    def set_launch(gi, g, o):
        state[gi][1:] = [g, o]

    def rk4_timed(steps):
        calls.append(steps)
        return np.array([(table[gi][state[gi][0]] + grid_gain[state[gi][1]]) * 4 * steps / 1e3 for gi in (0, 1)])

    b.groups = lambda: [{}, {}]
    b.set_tile_shape, b.set_launch, b.rk4_timed = set_tile_shape, set_launch, rk4_timed
    choices, advanced = b.tune_groups(steps=2, grids=(100, 200), orders=(0, 1), margin=0.98)
    # group 0: shape 1 at 96 us beats 0.98 x 99 (the built-in best); shape 3 at 98.5 does not
    # group 1: shape 2 at 39 us beats 0.98 x 49
    assert [c[:3] for c in choices] == [(1, 200, 0), (2, 200, 0)]
    assert choices[0][3] == pytest.approx(96.0) and choices[1][3] == pytest.approx(39.0)
    assert state == {0: [1, 200, 0], 1: [2, 200, 0]}
    # 4 tile shapes with a kernel on some group x 2 grids x 2 tile orders, 2 steps each
    assert advanced == len(calls) * 2 == 4 * 4 * 2
    table[0][1] = 98.5  # 97.5 with the grid gain, above 0.98 x 99 = 97.02: the built-in shape stays
    choices, _ = b.tune_groups(steps=1, grids=(100, 200), orders=(0,), margin=0.98)
    assert choices[0][:3] == (0, 200, 0)
