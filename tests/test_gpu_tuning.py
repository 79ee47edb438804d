"""Launch-shape tuning never changes results: every grid size and tile order gives bitwise the
same state; a tuning table is applied at context creation."""
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
