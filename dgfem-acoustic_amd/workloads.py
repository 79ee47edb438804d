"""Workload definitions shared by bench.py, the autotuner, the surrogate room driver and the tests
(meshes, physics, initial states).

Time steps use a CFL number of 0.1 relative to the DG node spacing, dt = 0.1 h_min /
(c (2p+1)): with dt = 0.1 h/c the order-3 schemes are unstable under RK4 (2D limit about
0.06 h/c, 3D about 0.05 h/c for these meshes).

Sizes follow the benchmark configurations: a structured 2D square at order 3, a jittered
unit cube with a Kuhn tetrahedral split at order 3, the 20x15x8 m room with a block
obstacle at order 4, the 256-member ensemble of uneven 2D/3D simulations at orders 2-5,
plus small variants of each for parity tests.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import dgk

BASE_CONF = ("timeStart=0\ntimeEnd=1\ntimeStep=1\ntimeRate=1\nelementType=Lagrange\ntimeIntMethod=Runge-Kutta\n"
             "numThreads=1\nv0_x=0\nv0_y=0\nv0_z=0\nrho0=1.225\nc0=340\n")


def conf(extra=""):
    return BASE_CONF + extra


def cfl_dt(hmin, speed, order, cfl=0.1):
    return cfl * hmin / (speed * (2 * order + 1))


def gaussian(mesh, center, size, amp=1.0):
    x = mesh.node_xyz
    u = np.zeros((4, mesh.K * mesh.Np))
    r2 = ((x[:, 0] - center[0]) ** 2 + (x[:, 1] - center[1]) ** 2 + (x[:, 2] - center[2]) ** 2)
    u[0] = amp * np.exp(-r2 / size)
    return u


def square_2d(n=40, order=3):
    """[0,1]^2, n x n cells, 2 CCW triangles each, one reflecting group 'Wall'."""
    m = dgk.gen_rect(n, n, 1.0, 1.0, order, [1, 1, 1, 1], "1:Wall", conf("Wall=Reflecting\n"))
    phys = dict(c0=340.0, rho0=1.225, v0=(0.0, 0.0, 0.0))
    h = 1.0 / n
    dt = cfl_dt(h, 340.0, order)
    u0 = gaussian(m, (0.5, 0.5, 0.0), 0.005)
    return m, phys, dt, u0


def cube_3d(n=32, order=3, jitter=0.1, seed=2):
    """[0,1]^3, n^3 cells Kuhn-split (6n^3 tets), interior nodes jittered, all faces absorbing."""
    m = dgk.gen_box(n, n, n, 1.0, 1.0, 1.0, order, jitter, seed, [1] * 6, "1:Abs", conf("Abs=Absorbing\n"))
    phys = dict(c0=340.0, rho0=1.225, v0=(50.0, 0.0, 0.0))
    rng = np.random.Generator(np.random.MT19937(3))
    center = 0.3 + 0.4 * rng.random(3)
    h = 1.0 / n
    hmin = (1 - 2 * jitter) * h
    dt = cfl_dt(hmin, 340.0 + 50.0, order)
    u0 = gaussian(m, center, 2 * 0.04 ** 2)
    return m, phys, dt, u0


def room_3d(h=0.25, order=4, seed=4):
    """20x15x8 m room, Kuhn tets of size h, a 2x2x1 m block on the floor; walls reflecting, ceiling absorbing."""
    nx, ny, nz = int(round(20 / h)), int(round(15 / h)), int(round(8 / h))
    rng = np.random.Generator(np.random.MT19937(seed))
    bx, by = int(2 / h), int(2 / h)
    i0 = int(rng.integers(0, nx - bx))
    j0 = int(rng.integers(0, ny - by))
    block = ((i0, j0, 0), (i0 + bx, j0 + by, int(round(1 / h))))
    m = dgk.gen_box(nx, ny, nz, 20.0, 15.0, 8.0, order, 0.0, 0, [1, 1, 1, 1, 1, 2], "1:Walls;2:Ceiling",
                    conf("Walls=Reflecting\nCeiling=Absorbing\nc0=343\n"), block=block, block_tag=1)
    phys = dict(c0=343.0, rho0=1.225, v0=(0.0, 0.0, 0.0))
    sigma = 343.0 / (2 * np.pi * 500.0)
    center = (4.0 + 12.0 * rng.random(), 3.0 + 9.0 * rng.random(), 1.5 + 5.0 * rng.random())
    dt = cfl_dt(h, 343.0, order)
    u0 = gaussian(m, center, 2 * sigma ** 2)
    return m, phys, dt, u0


def ensemble_draw(i, seed0=7, n=256, kmin=5000, kmax=80000, orders=(2, 3, 4, 5)):
    """The random draws of ensemble member i, without building its mesh: (dim, K_target, order,
    sides L, wall groups, sigma, centre)."""
    rng = np.random.Generator(np.random.MT19937(seed0 + i))
    dim = 2 if i < n // 2 else 3
    K_target = int(rng.integers(kmin, kmax + 1))
    p = int(orders[int(rng.integers(len(orders)))])
    L = 1.0 + 2.0 * rng.random(dim)
    walls = [1 if r < 0.5 else 2 for r in rng.random(2 * dim)]
    sigma = (0.02 + 0.06 * rng.random()) * float(L.min())
    center = [3 * sigma + (L[k] - 6 * sigma) * rng.random() for k in range(dim)] + [0.0] * (3 - dim)
    return dim, K_target, p, L, walls, sigma, center


def ensemble_member(i, seed0=7, n=256, kmin=5000, kmax=80000, orders=(2, 3, 4, 5), k_target=None):
    """Member i of the uneven ensemble: the first half 2D rectangles (2 triangles per cell), the
    rest 3D boxes (Kuhn split), sides U[1,3], K ~ U[kmin,kmax], order ~ U{orders}, each wall
    reflecting or absorbing with probability 1/2, Gaussian pulse with sigma ~ U[0.02,0.08] times
    the shortest side, centred at least 3 sigma from every wall. Seeds 7+i (MT19937). k_target
    replaces the drawn element count (a smaller copy of the member: same shape, walls and pulse)."""
    dim, K_target, p, L, walls, sigma, center = ensemble_draw(i, seed0, n, kmin, kmax, orders)
    if k_target is not None:
        K_target = k_target
    names = "1:Wall;2:Abs"
    text = conf("Wall=Reflecting\nAbs=Absorbing\n")
    if dim == 2:
        h = np.sqrt(2.0 * L[0] * L[1] / K_target)
        nx, ny = max(1, int(round(L[0] / h))), max(1, int(round(L[1] / h)))
        m = dgk.gen_rect(nx, ny, float(L[0]), float(L[1]), p, walls, names, text)
        hmin = min(L[0] / nx, L[1] / ny)
    else:
        h = np.cbrt(6.0 * L[0] * L[1] * L[2] / K_target)
        nx, ny, nz = (max(1, int(round(L[k] / h))) for k in range(3))
        m = dgk.gen_box(nx, ny, nz, float(L[0]), float(L[1]), float(L[2]), p, 0.0, 0, walls, names, text)
        hmin = min(L[0] / nx, L[1] / ny, L[2] / nz)
    m.release()
    phys = dict(c0=340.0, rho0=1.225, v0=(0.0, 0.0, 0.0))
    dt = cfl_dt(hmin, 340.0, p)
    return m, phys, dt, center, sigma


def ensemble(n=256, seed0=7, kmin=5000, kmax=80000, orders=(2, 3, 4, 5), threads=16, members=None, k_target=None):
    """The uneven ensemble (see ensemble_member), generated on a thread pool (the generators
    release the GIL). Returns lists (meshes, phys, dts, pulses); ensemble_state(mesh, pulse)
    builds a member's initial state."""
    idx = list(range(n)) if members is None else list(members)
    with ThreadPoolExecutor(max_workers=threads) as ex:
        out = list(ex.map(lambda i: ensemble_member(i, seed0, n, kmin, kmax, orders, k_target), idx))
    return [o[0] for o in out], [o[1] for o in out], [o[2] for o in out], [(o[3], o[4]) for o in out]


def ensemble_state(mesh, pulse):
    center, sigma = pulse
    return gaussian(mesh, center, 2 * sigma ** 2)


def dof_count(meshes):
    """Counted unknowns: (dim+1) fields per node."""
    return sum((m.dim + 1) * m.K * m.Np for m in meshes)


def room_box(Lx, Ly, Lz, h, order, block=None, seed=4):
    """Shoebox room of Kuhn tets of size h: walls and floor reflecting, ceiling absorbing, c0 = 343;
    optional block obstacle ((i0,j0,k0), (i1,j1,k1)) in cells. Returns (mesh, phys, dt, sigma) with
    the 500 Hz-equivalent pulse width sigma = c0 / (2 pi 500)."""
    nx, ny, nz = int(round(Lx / h)), int(round(Ly / h)), int(round(Lz / h))
    m = dgk.gen_box(nx, ny, nz, Lx, Ly, Lz, order, 0.0, 0, [1, 1, 1, 1, 1, 2], "1:Walls;2:Ceiling",
                    conf("Walls=Reflecting\nCeiling=Absorbing\nc0=343\n"), block=block, block_tag=1)
    phys = dict(c0=343.0, rho0=1.225, v0=(0.0, 0.0, 0.0))
    return m, phys, cfl_dt(h, 343.0, order), 343.0 / (2 * np.pi * 500.0)
