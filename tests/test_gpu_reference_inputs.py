"""The reference's own example inputs (its P1 MSH 4.0 meshes and example configs, copied as
fixtures by tests/golden/fetch_reference_inputs.py) through the product path on the GPU,
against the CPU oracle: mesh ingestion, physical-group BCs from the config, the Gaussian
initial condition and the monopole source expansion, then RK4 steps at the config's time step."""
import os

import numpy as np
import pytest

import dgk
import oracle_binding as ob
from test_gpu_parity import rel

pytestmark = pytest.mark.gpu

D = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_inputs")


def initial_state(m, cfg):
    u = np.zeros((4, m.K * m.Np))
    x = m.node_xyz
    for ic in cfg["initConditions"]:  # [kind, x, y, z, size, amp]: p += amp exp(-|x - x0|^2 / size)
        x0 = np.array(ic[1:4])
        u[0] += ic[5] * np.exp(-((x - x0) ** 2).sum(axis=1) / ic[4])
    return u


@pytest.mark.parametrize("mesh,conf,nsteps", [("line.msh", "config.conf", 40), ("square.msh", "config.conf", 40),
                                              ("room.msh", "Room_2D.conf", 40), ("cube.msh", "config.conf", 20)])
def test_reference_example_inputs(mesh, conf, nsteps):
    text = open(os.path.join(D, conf)).read()
    m = dgk.read_mesh(os.path.join(D, mesh), text)
    cfg = dgk.parse_config(text)
    dt = cfg["timeStep"]
    ts = cfg["timeStart"] + dt * np.arange(nsteps)
    t = cfg["timeStart"]
    for i in range(nsteps):  # the driver's floating-point accumulation of t
        ts[i] = t
        t += dt
    u0 = initial_state(m, cfg)
    src = np.array(cfg["sources"], dtype=np.float64).reshape(-1, 9)
    c0, rho0, v0 = cfg["c0"], cfg["rho0"], cfg["v0"]
    o = ob.Oracle(m, c0, rho0, v0=v0)
    s = dgk.Solver(m, c0, rho0, v0=v0)
    if len(src):
        s.set_sources(src)
    s.set_state(u0)
    s.rk4(nsteps, dt, ts)
    ref = o.step(u0, nsteps, dt, t_steps=ts, sources=src if len(src) else None)
    err = rel(s.get_state(), ref)
    print(f"{mesh} ({m.dim}D P{m.order}, K={m.K}) with {conf}: {nsteps} RK4 steps, rel L2 {err:.2e}")
    assert np.linalg.norm(ref) > 0 and err <= 1e-12
