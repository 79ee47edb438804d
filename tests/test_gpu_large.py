"""Meshes beyond 32-bit device offsets and beyond 2^21 vertices.

The reference has no size cap (its face search, src/Mesh.cpp:107-128, is only slow). Here the
stage kernel addresses arrays below 4 GiB through 32-bit offsets from the buffer bases and larger
ones through 64-bit per-tile bases (the same arithmetic), and 3D face keys hold three 32-bit
vertex ids (96 bits, two stable radix passes).

* The 64-bit addressing form, forced on small meshes of every dimension and both precisions,
  gives bitwise the results of the 32-bit form and matches the oracle.
* A 3D order-2 box of 132^3 Kuhn cells (2.30 M vertices > 2^21, 13.8 M tets, FP64 state buffers of
  4.4 GB > 4 GiB) against the oracle after two RK4 steps, state and increment.
"""
import numpy as np
import pytest

import dgk
import oracle_binding as ob
from test_gpu_parity import mixed_mesh, rel
from workloads import conf

pytestmark = pytest.mark.gpu

FORCED = [(1, 3, dgk.FP64), (2, 2, dgk.FP32), (2, 3, dgk.FP64), (2, 5, dgk.FP64), (2, 3, dgk.FP32), (3, 1, dgk.FP64),
          (3, 3, dgk.FP64), (3, 4, dgk.FP32), (3, 4, dgk.FP64), (3, 5, dgk.FP64), (3, 5, dgk.FP32)]


@pytest.mark.parametrize("dim,p,prec", FORCED)
def test_addr64_form_equals_32bit_form_and_oracle(dim, p, prec):
    print(f"addr64 dim={dim} p={p} precision={prec}", flush=True)
    m = mixed_mesh(dim, p, 4 if dim == 3 else None)
    x = m.node_xyz
    u0 = np.zeros((4, m.K * m.Np))
    u0[0] = np.exp(-((x - x.mean(axis=0)) ** 2).sum(axis=1) / 0.05)
    out = []
    for a64 in (False, True):
        s = dgk.Solver(m, 340.0, 1.225, v0=(20.0, -5.0, 3.0), precision=prec, flux_sign=dgk.SIGN_STABLE,
                       addr64=a64)
        s.set_tile_shape(0)  # the same tile shape in both forms (other shapes exist below 4 GiB only)
        s.set_state(u0)
        s.rk4(30, 1e-6)
        out.append(s.get_state())
        s.close()
    assert np.array_equal(out[0], out[1])
    o = ob.Oracle(m, 340.0, 1.225, v0=(20.0, -5.0, 3.0), sign_mode=ob.SIGN_STABLE)
    ref = o.step(u0, 30, 1e-6)
    assert rel(out[1], ref) <= (1e-12 if prec == dgk.FP64 else 1e-5)


@pytest.mark.timeout(1200)
def test_box_beyond_2e21_vertices_and_4gib_buffers():
    n = 132
    print(f"large box: {n}^3 Kuhn cells, order 2, FP64", flush=True)
    m = dgk.gen_box(n, n, n, 1.0, 1.0, 1.0, 2, 0.1, 2, [1, 2, 1, 2, 1, 2], "1:Wall;2:Abs",
                    conf("Wall=Reflecting\nAbs=Absorbing\n"))
    m.release()
    assert m.K == 6 * n ** 3
    assert len(np.unique(m.node_tags.reshape(m.K, m.Np)[:, :4])) == (n + 1) ** 3 > 2 ** 21
    # the pulse sits near the top of the box, where the last elements (those whose state lies
    # beyond the first 4 GiB of each buffer) are
    x = m.node_xyz
    u0 = np.zeros((4, m.K * m.Np))
    u0[0] = np.exp(-((x - [0.5, 0.5, 0.95]) ** 2).sum(axis=1) / 0.01)
    u0[1] = 0.01 * np.exp(-((x - [0.45, 0.5, 0.97]) ** 2).sum(axis=1) / 0.02)
    dt = 1e-6
    s = dgk.Solver(m, 340.0, 1.225, v0=(20.0, 0.0, 0.0), flux_sign=dgk.SIGN_STABLE)
    assert 4 * m.K * m.Np * 8 >= 2 ** 32  # one state buffer
    s.set_state(u0)
    s.rk4(2, dt)
    g = s.get_state()
    print(f"  device bytes {s.info.device_bytes / 2 ** 30:.1f} GiB", flush=True)
    s.close()
    del s
    o = ob.Oracle(m, 340.0, 1.225, v0=(20.0, 0.0, 0.0), sign_mode=ob.SIGN_STABLE, factorized_only=True)
    ref = o.step(u0, 2, dt)
    err = rel(g, ref)
    kb = 2 ** 32 // (m.Np * 4 * 8)  # first element whose interleaved state starts beyond 4 GiB
    assert 0 < kb < m.K
    parts = {}
    for name, sl in (("below 4 GiB", slice(0, kb * m.Np)), ("beyond 4 GiB", slice(kb * m.Np, m.K * m.Np))):
        d, inc = (g - ref)[:, sl], (ref - u0)[:, sl]
        parts[name] = float(np.linalg.norm(d) / np.linalg.norm(inc))
    print(f"  two RK4 steps, rel-L2 vs oracle {err:.2e}; of the increment {parts}", flush=True)
    assert err <= 1e-12
    assert max(parts.values()) <= 1e-10
