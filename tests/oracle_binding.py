"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: imported by tests/, the smoke entry point and bench.py's
CPU-baseline leg only, as the checker.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

LITERAL, FACTORIZED, NESTED_FACES = 0, 1, 0x10
SIGN_REFERENCE, SIGN_STABLE = 0, 1

_lib = None


def load(path):
    """Load an oracle library build (the shipped one, or a mutant built by the pin tests)."""
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
    L = ctypes.CDLL(path)
    _declare(L)
    return L


def lib():
    global _lib
    if _lib is None:
        _lib = load(LIB_PATH)
    return _lib


def _declare(L):
    P = ctypes.POINTER
    L.orc_create.restype = ctypes.c_void_p
    L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, P(ctypes.c_double), P(ctypes.c_int64),
                             ctypes.c_int, P(ctypes.c_int), P(P(ctypes.c_int64)), P(ctypes.c_int64),
                             ctypes.c_double, ctypes.c_double, P(ctypes.c_double), ctypes.c_int, ctypes.c_int,
                             ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    L.orc_destroy.argtypes = [ctypes.c_void_p]
    L.orc_sizes.argtypes = [ctypes.c_void_p, P(ctypes.c_int64)]
    L.orc_connectivity.argtypes = [ctypes.c_void_p] + [P(ctypes.c_int32)] * 5
    L.orc_face_normals.argtypes = [ctypes.c_void_p, P(ctypes.c_double)]
    L.orc_mass_inverse.argtypes = [ctypes.c_void_p, P(ctypes.c_double)]
    L.orc_rhs.argtypes = [ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_double), ctypes.c_double, ctypes.c_int]
    L.orc_step.argtypes = [ctypes.c_void_p, P(ctypes.c_double), ctypes.c_int, ctypes.c_double, P(ctypes.c_double),
                           ctypes.c_int, P(ctypes.c_double), ctypes.c_int, ctypes.c_int]
    L.orc_ref_matrices.argtypes = [ctypes.c_int, ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double),
                                   P(ctypes.c_double)]
    L.orc_ref_nodes.argtypes = [ctypes.c_int, ctypes.c_int, P(ctypes.c_double)]


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


class Oracle:
    """CPU restatement of the reference solver on one mesh (see oracle/oracle.h)."""

    def __init__(self, mesh, c0, rho0, v0=(0, 0, 0), sign_mode=SIGN_REFERENCE, nthreads=0, literal_setup=False,
                 factorized_only=False, library=None):
        L = self.L = library or lib()
        self.mesh = mesh
        self.dim, self.order, self.K, self.Np = mesh.dim, mesh.order, mesh.K, mesh.Np
        xyz = np.ascontiguousarray(mesh.node_xyz, dtype=np.float64)
        tags = np.ascontiguousarray(mesh.node_tags, dtype=np.int64)
        groups = mesh.bc_groups  # list of (tag, reflecting, np.int64 nodes)
        ng = len(groups)
        refl = np.array([g[1] for g in groups] or [0], dtype=np.int32)
        arrs = [np.ascontiguousarray(g[2], dtype=np.int64) for g in groups]
        ptrs = (ctypes.POINTER(ctypes.c_int64) * max(ng, 1))(*[_p(a, ctypes.c_int64) for a in arrs])
        sizes = np.array([len(a) for a in arrs] or [0], dtype=np.int64)
        v0a = np.array(v0, dtype=np.float64)
        err = ctypes.create_string_buffer(512)
        self._keep = (xyz, tags, arrs, ptrs, refl, sizes, v0a)
        h = L.orc_create(self.dim, self.order, self.K, _p(xyz), _p(tags, ctypes.c_int64), ng, _p(refl, ctypes.c_int32),
                         ptrs, _p(sizes, ctypes.c_int64), float(c0), float(rho0), _p(v0a), int(sign_mode),
                         int(nthreads), int(literal_setup) | (2 if factorized_only else 0), err, 512)
        if not h:
            raise RuntimeError("oracle: " + err.value.decode())
        self.h = ctypes.c_void_p(h)
        s = np.zeros(8, dtype=np.int64)
        L.orc_sizes(self.h, _p(s, ctypes.c_int64))
        self.Nf, self.Nfp, self.F, self.Fb, self.G, self.Gf, self.fc = (int(x) for x in s[1:])

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def connectivity(self):
        K, Nf, F, Nfp = self.K, self.Nf, self.F, self.Nfp
        elF = np.zeros(K * Nf, np.int32)
        elO = np.zeros(K * Nf, np.int32)
        nbr = np.zeros(2 * F, np.int32)
        f2e = np.zeros(F * Nfp * 2, np.int32)
        bc = np.zeros(F, np.int32)
        self.L.orc_connectivity(self.h, *[_p(a, ctypes.c_int32) for a in (elF, elO, nbr, f2e, bc)])
        return dict(elFIds=elF.reshape(K, Nf), elFOrientation=elO.reshape(K, Nf), fNbr=nbr.reshape(F, 2),
                    fNToElNIds=f2e.reshape(F, Nfp, 2), fBC=bc)

    def face_normals(self):
        out = np.zeros((self.F, 3))
        self.L.orc_face_normals(self.h, _p(out))
        return out

    def mass_inverse(self):
        out = np.zeros((self.K, self.Np, self.Np))
        self.L.orc_mass_inverse(self.h, _p(out))
        return out

    def rhs(self, u, dt, mode=FACTORIZED):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros_like(u)
        if self.L.orc_rhs(self.h, _p(u), _p(out), float(dt), int(mode)):
            raise RuntimeError("oracle: literal evaluation needs the literal setup")
        return out

    def step(self, u, nsteps, dt, mode=FACTORIZED, t_steps=None, sources=None, euler=False):
        u = np.array(u, dtype=np.float64, copy=True, order="C")
        ts = None if t_steps is None else np.ascontiguousarray(t_steps, dtype=np.float64)
        src = np.zeros((0, 9)) if sources is None else np.ascontiguousarray(sources, dtype=np.float64).reshape(-1, 9)
        rc = self.L.orc_step(self.h, _p(u), int(nsteps), float(dt), None if ts is None else _p(ts), len(src),
                            _p(src), int(mode), int(euler))
        if rc:
            raise RuntimeError("oracle: literal evaluation needs the literal setup")
        return u


def ref_matrices(dim, order):
    from math import comb
    Np = comb(order + dim, dim)
    Nfp = 1 if dim == 1 else comb(order + dim - 1, dim - 1)
    M = np.zeros((Np, Np))
    S = np.zeros((dim, Np, Np))
    Mf = np.zeros((Nfp, Nfp))
    rc = lib().orc_ref_matrices(dim, order, _p(M), _p(S), _p(Mf))
    assert rc == 0
    return M, S, Mf


def ref_nodes(dim, order):
    from math import comb
    Np = comb(order + dim, dim)
    out = np.zeros((Np, 3))
    lib().orc_ref_nodes(dim, order, _p(out))
    return out
