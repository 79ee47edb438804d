"""Launch-shape tuning never changes results: every grid size and tile order gives bitwise the
same state; every tile shape agrees with the built-in one (bitwise, or to rounding where the row
split sums in another order); a tuning table is applied at context creation."""
import os

import numpy as np
import pytest

import dgk
from test_gpu_parity import mixed_mesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,p,prec", [(2, 3, dgk.FP64), (3, 3, dgk.FP32), (3, 5, dgk.FP32)])
def test_launch_shape_is_bitwise_neutral(dim, p, prec):
    m = mixed_mesh(dim, p, 5)
    u = np.random.default_rng(2).standard_normal((4, m.K * m.Np))
    out = []
    for g, o in ((100, 0), (50, 1), (400, 0), (200, 1), (1, 0)):
        s = dgk.Solver(m, 340.0, 1.225, precision=prec, flux_sign=dgk.SIGN_STABLE)
        s.set_launch(g, o)
        s.set_state(u)
        s.rk4(30, 1e-6)
        out.append(s.get_state())
        s.close()
    for x in out[1:]:
        assert np.array_equal(x, out[0])


def test_tuning_table_is_loaded(tmp_path):
    L = dgk.device()
    f = tmp_path / "t.txt"
    f.write_text("# test\n2 3 0 2 50 1 1.0\n")
    assert L.dgk_load_tuning(str(f).encode()) == 0
    bad = tmp_path / "bad.txt"
    bad.write_text("2 3 0 7 50 1\n")
    assert L.dgk_load_tuning(str(bad).encode()) == dgk.DGK_ERR_INVALID
    m = mixed_mesh(2, 3)
    u = np.random.default_rng(3).standard_normal((4, m.K * m.Np))
    s = dgk.Solver(m, 340.0, 1.225)
    s.set_state(u)
    s.rk4(10, 1e-6)
    tuned = s.get_state()
    # restore the shipped table (or none)
    empty = tmp_path / "empty.txt"
    empty.write_text("")
    assert L.dgk_load_tuning((dgk.TUNING_FILE if os.path.exists(dgk.TUNING_FILE) else str(empty)).encode()) == 0
    s2 = dgk.Solver(m, 340.0, 1.225)
    s2.set_state(u)
    s2.rk4(10, 1e-6)
    assert np.array_equal(tuned, s2.get_state())


SHAPE_CASES = [(d, p, prec) for d in (2, 3) for p in (2, 3, 4, 5) for prec in (dgk.FP64, dgk.FP32)]


def _row_split(shape, dim, p, prec):
    """Whether the MFMA kernel of this tile shape splits operator rows across waves (the built-in
    shape does for FP64 3D order 5)."""
    return shape in (3, 5) or (shape == 0 and prec == dgk.FP64 and dim == 3 and p == 5)


def _waves_per_simd_budget(shape, dim, p, prec):
    """Waves per SIMD the tile shape's registers are sized for (shapes 6-8: 3-5, 9-10: 4-5, 11: 3;
    the built-in shape of FP32 3D order 4 is shape 6)."""
    if shape == 0:
        return 3 if (prec == dgk.FP32 and dim == 3 and p == 4) else 1
    return shape - 3 if shape in (6, 7, 8) else shape - 5 if shape in (9, 10) else 3 if shape == 11 else 1


def _run(m, u, prec, shape=0, steps=30):
    """RK4 steps on tile shape `shape` (0: the built-in one, whatever the tuning table says; None:
    the shape the context was created with)."""
    s = dgk.Solver(m, 340.0, 1.225, v0=(30.0, -10.0, 5.0), precision=prec, flux_sign=dgk.SIGN_STABLE)
    try:
        if shape is not None:
            s.set_tile_shape(shape)
        s.set_state(u)
        s.rk4(steps, 1e-6)
        return s.get_state()
    finally:
        s.close()


@pytest.mark.parametrize("dim,p,prec", SHAPE_CASES)
def test_tile_shapes_agree_with_the_built_in_shape(dim, p, prec):
    """Every tile shape with a kernel for the case computes the same discrete operator: bitwise the
    built-in shape's state when both split elements across waves (the same k-step order per output)
    and are compiled for the same waves per SIMD; otherwise within rounding (FP64 1e-12, FP32 1e-5
    relative L2 after 30 steps): the row split sums the k-steps in two partial accumulators, and
    another register budget may fuse a product into a multiply-add differently (measured: shape 8
    differs from the built-in shape by 1e-16 FP64 and 1e-8 FP32 relative after 30 steps). A shape
    without a kernel is refused with DGK_ERR_UNSUPPORTED and leaves the context on its shape."""
    print(f"case dim={dim} p={p} prec={prec}", flush=True)
    m = mixed_mesh(dim, p, 9)
    u = np.random.default_rng(5).standard_normal((4, m.K * m.Np))
    if dim == 2:
        u[3] = 0  # the 2D form with its 3 live fields, which the other tile shapes are built for
    ref = _run(m, u, prec)
    ran = []
    for sh in dgk.TILE_SHAPES[1:]:
        try:
            x = _run(m, u, prec, sh)
        except dgk.DGKError as e:
            assert e.status == dgk.DGK_ERR_UNSUPPORTED, (sh, e)
            continue
        ran.append(sh)
        if (_row_split(sh, dim, p, prec) == _row_split(0, dim, p, prec) and
                _waves_per_simd_budget(sh, dim, p, prec) == _waves_per_simd_budget(0, dim, p, prec)):
            assert np.array_equal(x, ref), sh
        else:
            err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
            print(f"  shape {sh}: rel L2 {err:.2e} against the built-in shape", flush=True)
            assert err <= (1e-12 if prec == dgk.FP64 else 1e-5), (sh, err)
    # the 4-wave element-split shapes exist for every case here; the row splits (4 and 8 waves) only
    # where the operator is whole 16-row blocks (order 5 in 3D)
    assert {1, 4} <= set(ran), ran
    assert (3 in ran) == (5 in ran) == (dim == 3 and p == 5), ran


def _batch(m2, u2, m3, u3, shapes=None):
    b = dgk.Batch([m2, m3], [dict(c0=340.0, rho0=1.225, v0=(30.0, -10.0, 5.0))] * 2, [1e-6, 1e-6],
                  precision=dgk.FP32, flux_sign=dgk.SIGN_STABLE)
    try:
        g = {(x["dim"], x["order"]): i for i, x in enumerate(b.groups())}
        for key, sh in (shapes or {}).items():
            b.set_tile_shape(g[key], sh)
        b.set_state(0, u2)
        b.set_state(1, u3)
        b.rk4(30)
        return b.get_state(0), b.get_state(1)
    finally:
        b.close()


def test_tile_shape_fallback_and_batch_groups(tmp_path):
    """A 2D context whose state gains a nonzero vz runs the built-in shape (bitwise the default); a
    batch group takes a tile shape of its own (refused where it has no such kernel); the tuning
    table's tile-shape column is applied at creation where the context has such a kernel and
    ignored where it has none."""
    m2 = mixed_mesh(2, 3, 9)
    u2 = np.random.default_rng(6).standard_normal((4, m2.K * m2.Np))  # vz nonzero: 4 live fields
    assert np.array_equal(_run(m2, u2, dgk.FP32, 2), _run(m2, u2, dgk.FP32))
    m3 = mixed_mesh(3, 5, 9)
    u3 = np.random.default_rng(7).standard_normal((4, m3.K * m3.Np))
    u2[3] = 0
    ref2, ref3 = _run(m2, u2, dgk.FP32), _run(m3, u3, dgk.FP32)
    row3 = _run(m3, u3, dgk.FP32, 3)
    assert not np.array_equal(row3, ref3)  # the row split sums in another order
    for key, sh, status in (((2, 3), 3, dgk.DGK_ERR_UNSUPPORTED), ((3, 5), len(dgk.TILE_SHAPES), dgk.DGK_ERR_INVALID)):
        with pytest.raises(dgk.DGKError) as e:
            _batch(m2, u2, m3, u3, {key: sh})
        assert e.value.status == status
    b2, b3 = _batch(m2, u2, m3, u3, {(2, 3): 0, (3, 5): 0})  # the built-in shapes
    s2, s3 = _batch(m2, u2, m3, u3, {(2, 3): 2, (3, 5): 3})
    assert np.array_equal(s2, b2)  # 8 waves x 32 elements: same k-step order
    err = np.linalg.norm(s3 - b3) / np.linalg.norm(b3)
    assert 0 < err <= 1e-5, err
    L = dgk.device()
    f = tmp_path / "t.txt"
    f.write_text("3 5 1 2 100 1 3 1.0\n2 3 1 2 100 1 3 1.0\n")  # 2D order 3 has no row split
    assert L.dgk_load_tuning(str(f).encode()) == 0
    try:
        t3, t2 = _run(m3, u3, dgk.FP32, None), _run(m2, u2, dgk.FP32, None)
    finally:
        empty = tmp_path / "empty.txt"
        empty.write_text("")
        assert L.dgk_load_tuning((dgk.TUNING_FILE if os.path.exists(dgk.TUNING_FILE) else str(empty)).encode()) == 0
    assert np.array_equal(t3, row3)
    assert np.array_equal(t2, ref2)


def test_batch_group_tuning_keeps_results():
    """Batch.tune_groups times every (tile shape, grid, tile order) candidate on every group with a
    kernel of that shape and keeps each group's fastest: the choices are valid shapes of the groups,
    the members advance exactly the timed steps, and a run after tuning agrees with an untuned batch
    advanced by the same steps to FP32 rounding."""
    m2, m3 = mixed_mesh(2, 3, 9), mixed_mesh(3, 4, 9)
    u2 = np.random.default_rng(8).standard_normal((4, m2.K * m2.Np))
    u3 = np.random.default_rng(9).standard_normal((4, m3.K * m3.Np))
    u2[3] = 0
    out = []
    for tune in (False, True):
        b = dgk.Batch([m2, m3], [dict(c0=340.0, rho0=1.225, v0=(30.0, -10.0, 5.0))] * 2, [1e-6, 1e-6],
                      precision=dgk.FP32, flux_sign=dgk.SIGN_STABLE)
        try:
            b.set_state(0, u2)
            b.set_state(1, u3)
            if tune:
                choices, advanced = b.tune_groups(steps=1, grids=(100, 200), orders=(0, 1))
                assert len(choices) == 2
                for sh, g, o, us in choices:
                    assert sh in dgk.TILE_SHAPES and g in (100, 200) and o in (0, 1) and us > 0
                # 2 launch grids x 2 tile orders per tile shape that some group has a kernel of
                assert advanced % 4 == 0 and advanced >= 4 * 8, advanced
            else:
                advanced = None
            out.append((b, advanced))
        except BaseException:
            b.close()
            raise
    (b0, _), (b1, advanced) = out
    try:
        b0.rk4(advanced)
        b0.rk4(20)
        b1.rk4(20)
        for i in (0, 1):
            x, y = b1.get_state(i), b0.get_state(i)
            err = np.linalg.norm(x - y) / np.linalg.norm(y)
            print(f"  member {i}: tuned vs untuned rel L2 {err:.2e}", flush=True)
            assert err <= 1e-5
    finally:
        b0.close()
        b1.close()
