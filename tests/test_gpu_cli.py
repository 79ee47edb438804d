"""The `dgalerkin mesh.msh config.conf` driver end to end on the GPU: MSH in, config keys, the
reference's time loop and save schedule, ElementNodeData views and receiver series out, checked
against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

import dgk
import oracle_binding as ob
from interp import sample
from test_gpu_parity import mixed_mesh, rel

pytestmark = pytest.mark.gpu

EXE = os.path.join(dgk.PKG_DIR, "dgalerkin")


def read_views(path, name):
    """[(step, time, values[K*Np*ncomp])] of the ElementNodeData views called `name`."""
    lines = open(path).read().splitlines()
    out, i = [], 0
    while i < len(lines):
        if lines[i] == "$ElementNodeData" and lines[i + 2] == f'"{name}"':
            t = float(lines[i + 4])
            step, ncomp, ne = int(lines[i + 6]), int(lines[i + 7]), int(lines[i + 8])
            vals = []
            for r in lines[i + 9:i + 9 + ne]:
                f = r.split()
                vals.extend(float(v) for v in f[2:])
            out.append((step, t, np.array(vals)))
            i += 9 + ne
        else:
            i += 1
    return out


def config(out, dt, n, precision="fp64"):
    return "\n".join([
        "timeStart = 0", f"timeEnd = {(n - 0.5) * dt!r}", f"timeStep = {dt!r}", f"timeRate = {10 * dt!r}",
        f"saveFile = {out}", "elementType = Lagrange", "timeIntMethod = Runge-Kutta", "numThreads = 1",
        "v0_x = 0", "v0_y = 0", "v0_z = 0", "rho0 = 1.225", "c0 = 340", "Wall = Reflecting", "Abs = Absorbing",
        "initialCondtition1 = gaussian, 0.5, 0.4, 0, 0.01, 1", "receiver1 = 0.31, 0.42, 0",
        "receiver2 = 0.7, 0.23, 0", f"gpu.precision = {precision}", "# a comment line", ""])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_cli_views_and_receivers_match_oracle(tmp_path, precision):
    if not os.path.exists(EXE):
        pytest.skip("dgalerkin not built")
    m = mixed_mesh(2, 3)
    msh, conf_path, out = tmp_path / "mesh.msh", tmp_path / "run.conf", tmp_path / "out.msh"
    m.write(str(msh))
    dt, n = 2e-6, 35
    conf_path.write_text(config(str(out), dt, n, precision))
    r = subprocess.run([EXE, str(msh), str(conf_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    views = read_views(str(out), "Pressure")
    assert views[0][0] == 0 and len(views) >= 3
    o = ob.Oracle(m, 340.0, 1.225)  # reference sign, the config default
    x = m.node_xyz
    u0 = np.zeros((4, m.K * m.Np))
    u0[0] = np.exp(-((x - np.array([0.5, 0.4, 0.0])) ** 2).sum(axis=1) / 0.01)
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert rel(views[0][2], u0[0]) <= (1e-15 if precision == "fp64" else 1e-7)
    step, t, p = views[-1]
    assert abs(t - step * dt) <= 1e-9 * dt * step
    assert rel(p, o.step(u0, step, dt)[0]) <= tol
    # receivers: one row per step, pressure interpolated at the points
    rec = np.loadtxt(str(out) + ".receivers.csv", delimiter=",", skiprows=1)
    assert rec.shape == (n, 3)
    pts = np.array([[0.31, 0.42, 0.0], [0.7, 0.23, 0.0]])
    u, ref = u0, []
    for _ in range(n):
        u = o.step(u, 1, dt)
        ref.append(sample(m, u, pts))
    assert rel(rec[:, 1:], np.array(ref)) <= tol
