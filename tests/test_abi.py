"""The C-ABI libraries load and export every symbol their headers declare; argument
validation that happens before any device call is exercised without a GPU."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import dgk
from workloads import conf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dg[khs]_[a-z0-9_]+)\s*\(", text)))


def ensure_device_lib():
    if not os.path.exists(dgk.DEV_LIB):
        subprocess.run([sys.executable, os.path.join(dgk.PKG_DIR, "build.py"), "device"], check=True)


def test_libdgk_exports_header():
    ensure_device_lib()
    lib = ctypes.CDLL(dgk.DEV_LIB)
    names = declared("dgk.h")
    assert "dgk_create" in names and "dgk_rk4" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(dgk.DEV_SYMBOLS)


def test_libdgk_exports_surrogate_header():
    import surrogate

    ensure_device_lib()
    lib = ctypes.CDLL(dgk.DEV_LIB)
    names = declared("dgk_surrogate.h")
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(surrogate.SYMBOLS)


def test_surrogate_rejects_bad_shapes_before_device_use():
    import surrogate

    ensure_device_lib()
    with pytest.raises(RuntimeError, match="multiples of 64"):
        surrogate.Net(width=100)


def test_libdgkhost_exports_header():
    lib = ctypes.CDLL(dgk.HOST_LIB)
    for n in declared("dgk_host.h"):
        if n == "dgh_node_bc_resolve":
            continue
        assert hasattr(lib, n), n


def _create(dim, order, K, xyz, tags, bc=None):
    ensure_device_lib()
    L = dgk.device()
    xyz = np.ascontiguousarray(xyz, np.float64)
    tags = np.ascontiguousarray(tags, np.int64)
    m = dgk._Mesh(dim, order, K, dgk._p(xyz), dgk._p(tags, ctypes.c_int64), None)
    ph = dgk._Phys(340.0, 1.225, (ctypes.c_double * 3)(0, 0, 0))
    op = dgk._Opts(0, 0, 0, 0)
    h = ctypes.c_void_p()
    st = L.dgk_create(ctypes.byref(m), ctypes.byref(ph), ctypes.byref(op), ctypes.byref(h))
    return st, L.dgk_last_error(None).decode(), h


def test_invalid_arguments_fail_before_device_use():
    st, msg, h = _create(2, 6, 1, np.zeros(3 * 28), np.arange(28))
    assert st == 2 and "order" in msg and not h
    st, msg, h = _create(4, 1, 1, np.zeros(12), np.arange(4))
    assert st == 2 and not h
    st, msg, h = _create(2, 1, 0, np.zeros(9), np.arange(3))
    assert st == 1 and not h


def test_curved_element_is_rejected():
    m = dgk.gen_rect(1, 1, 1.0, 1.0, 2, [1, 1, 1, 1], "1:Wall", conf("Wall=Reflecting\n"))
    xyz = m.node_xyz.copy()
    xyz[3, 1] += 0.05  # bend the first edge node of element 0
    st, msg, h = _create(2, 2, m.K, xyz, m.node_tags)
    assert st == 2 and "curved" in msg and not h


def test_shipped_tuning_table_covers_every_shape_and_loads(tmp_path):
    """tuning.txt has one valid row per (dim, order, precision) and the library accepts it; a row
    naming an unknown variant or an out-of-range launch shape is rejected (no device use)."""
    ensure_device_lib()
    rows = [l.split() for l in open(dgk.TUNING_FILE) if l.strip() and not l.startswith("#")]
    keys = sorted((int(r[0]), int(r[1]), int(r[2])) for r in rows)
    assert keys == sorted((d, p, q) for d in (1, 2, 3) for p in range(1, 6) for q in (dgk.FP64, dgk.FP32))
    for r in rows:
        assert int(r[3]) in (dgk.VARIANT_LDS, dgk.VARIANT_MFMA) and 1 <= int(r[4]) <= 10000 and int(r[5]) in (0, 1)
        assert len(r) == 8 and int(r[6]) in dgk.TILE_SHAPES  # dim order precision variant grid order shape us
        if int(r[2]) == dgk.FP32:
            assert int(r[3]) == dgk.VARIANT_MFMA  # the compensated FP32 form runs on the MFMA kernel
    L = dgk.device()
    assert L.dgk_load_tuning(dgk.TUNING_FILE.encode()) == 0
    for bad in ("3 3 1 9 100 0\n", "3 3 1 2 0 0\n", "3 3 1 2 100 2\n", "4 3 1 2 100 0\n", "3 3 1\n",
                f"3 3 1 2 100 0 {len(dgk.TILE_SHAPES)} 210.5\n", "3 3 1 2 100 0 -1 210.5\n", "3 3 1 2 100 0 1.5 210.5\n",
                "3 3 1 2 100 0 1 210.5 7\n"):
        f = tmp_path / "bad.txt"
        f.write_text(bad)
        assert L.dgk_load_tuning(str(f).encode()) == dgk.DGK_ERR_INVALID, bad
    # rows with the tile-shape column, and rows of tables written before it (stage_us only)
    for good in ("3 5 1 2 100 1 3 297.08\n", "3 5 1 2 100 1 297.08\n", "3 5 1 2 100 1\n"):
        f = tmp_path / "good.txt"
        f.write_text(good)
        assert L.dgk_load_tuning(str(f).encode()) == 0, good
    assert L.dgk_load_tuning(dgk.TUNING_FILE.encode()) == 0


def test_stage_addressing_switches_at_4_gib_per_array():
    """The stage kernel's 64-bit per-tile addressing starts exactly where one of its arrays reaches
    4 GiB: the state (Kpad*Np*fields*word) for 3D order 2, the J^-1 array (Kpad*9*word) for 1D
    order 1, whose state is smaller; element counts are padded to multiples of 32. Beyond 2^31
    element faces a context is refused (no device use)."""
    ensure_device_lib()
    lim = 1 << 32
    # 3D P2 FP64: 10 nodes x 4 fields x 8 B = 320 B per element of state
    k = lim // 320 // 32 * 32  # the largest padded count below the limit
    assert not dgk.stage_addressing(3, 2, dgk.FP64, k)
    assert dgk.stage_addressing(3, 2, dgk.FP64, k + 1)
    assert not dgk.stage_addressing(3, 2, dgk.FP32, k + 1)  # half the bytes
    # 1D P1 FP64: state 2 x 2 x 8 = 32 B per element, J^-1 9 x 8 = 72 B: J^-1 decides
    k1 = lim // 72 // 32 * 32
    assert 72 * k1 < lim <= 72 * (k1 + 32)
    assert not dgk.stage_addressing(1, 1, dgk.FP64, k1)
    assert dgk.stage_addressing(1, 1, dgk.FP64, k1 + 1)
    # 2D P5 FP32: 21 nodes x 3 fields x 4 B = 252 B of state against 4 x 3 faces x 4 B of face geometry
    k2 = -(-lim // 252)
    assert dgk.stage_addressing(2, 5, dgk.FP32, k2) and not dgk.stage_addressing(2, 5, dgk.FP32, k2 - 32)
    with pytest.raises(dgk.DGKError) as e:
        dgk.stage_addressing(3, 1, dgk.FP64, (1 << 31) // 4)
    assert e.value.status == dgk.DGK_ERR_UNSUPPORTED
    with pytest.raises(dgk.DGKError) as e:  # no overflow in the face count of absurd sizes
        dgk.stage_addressing(1, 1, dgk.FP64, (1 << 62) + 1)
    assert e.value.status == dgk.DGK_ERR_UNSUPPORTED
    with pytest.raises(dgk.DGKError) as e:
        dgk.stage_addressing(4, 1, dgk.FP64, 10)
    assert e.value.status == dgk.DGK_ERR_INVALID
