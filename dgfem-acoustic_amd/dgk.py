"""Python binding of the dgalerkin MI355X libraries (ctypes, no torch types).

libdgkhost.so : MSH ingestion, config semantics, mesh generators (include/dgk_host.h)
libdgk.so     : the DG time-stepping hot path on gfx950 (include/dgk.h)

The product path is the HIP library. If libdgk.so is missing or no GPU is
usable, the solver entry points raise; there is no CPU fallback.
"""
import ctypes
import json
import os
import time

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(PKG_DIR, "libdgkhost.so")
DEV_LIB = os.environ.get("DGK_LIBDGK") or os.path.join(PKG_DIR, "libdgk.so")
# launch settings per (dim, order, precision) written by autotune.py; loaded with the library
TUNING_FILE = os.environ.get("DGK_TUNING") or os.path.join(PKG_DIR, "tuning.txt")

_host = None
_dev = None
P = ctypes.POINTER


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(P(t))


def host():
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(HOST_LIB)
        L.dgh_last_error.restype = ctypes.c_char_p
        for fn in ("dgh_read", "dgh_gen_rect", "dgh_gen_box", "dgh_gen_line"):
            getattr(L, fn).restype = ctypes.c_void_p
        L.dgh_read.argtypes = [ctypes.c_char_p]
        L.dgh_gen_rect.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   P(ctypes.c_int), ctypes.c_char_p]
        L.dgh_gen_box.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [ctypes.c_int, ctypes.c_double,
                                  ctypes.c_uint64, P(ctypes.c_int), P(ctypes.c_int), ctypes.c_int, P(ctypes.c_int),
                                  ctypes.c_char_p]
        L.dgh_gen_line.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, P(ctypes.c_int), ctypes.c_char_p]
        L.dgh_free.argtypes = [ctypes.c_void_p]
        L.dgh_write.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.dgh_top_info.argtypes = [ctypes.c_void_p, P(ctypes.c_int64)]
        L.dgh_top_arrays.argtypes = [ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_int64), P(ctypes.c_int64)]
        L.dgh_config_json.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.dgh_bc_groups.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, P(ctypes.c_int), P(ctypes.c_int),
                                    P(ctypes.c_int), P(ctypes.c_int64)]
        L.dgh_group_nodes.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_int64)]
        L.dgh_node_bc.argtypes = [ctypes.c_void_p, ctypes.c_char_p, P(ctypes.c_uint8)]
        L.dgh_ref_element.argtypes = [ctypes.c_int, ctypes.c_int] + [P(ctypes.c_double)] * 6 + [P(ctypes.c_int)]
        _host = L
    return _host


def _check_handle(h):
    if not h:
        raise RuntimeError(host().dgh_last_error().decode())
    return ctypes.c_void_p(h)


def ref_element(dim, p):
    """Product reference-element operators (csrc/refelem.hpp) as numpy arrays."""
    from math import comb
    Np = comb(p + dim, dim)
    Nf = dim + 1
    Nfp = 1 if dim == 1 else comb(p + dim - 1, dim - 1)
    out = dict(nodes=np.zeros((Np, 3)), M=np.zeros((Np, Np)), S=np.zeros((dim, Np, Np)), Mf=np.zeros((Nfp, Nfp)),
               W=np.zeros((dim, Np, Np)), LIFT=np.zeros((Np, Nf * Nfp)), fmask=np.zeros((Nf, Nfp), np.int32))
    rc = host().dgh_ref_element(dim, p, *[_p(out[k]) for k in ("nodes", "M", "S", "Mf", "W", "LIFT")],
                                _p(out["fmask"], ctypes.c_int))
    if rc:
        raise ValueError(host().dgh_last_error().decode())
    return out


def parse_config(text):
    buf = ctypes.create_string_buffer(1 << 20)
    rc = host().dgh_config_json(text.encode(), buf, len(buf))
    if rc:
        raise ValueError(host().dgh_last_error().decode())
    return json.loads(buf.value.decode())


class Mesh:
    """The element block the solver works on plus its boundary groups (host arrays)."""

    def __init__(self, handle, config_text):
        L = host()
        self.h = handle
        info = np.zeros(4, np.int64)
        L.dgh_top_info(self.h, _p(info, ctypes.c_int64))
        self.dim, self.order, self.K, self.Np = (int(x) for x in info)
        self.node_xyz = np.zeros((self.K * self.Np, 3))
        self.node_tags = np.zeros(self.K * self.Np, np.int64)
        self.elem_tags = np.zeros(self.K, np.int64)
        L.dgh_top_arrays(self.h, _p(self.node_xyz), _p(self.node_tags, ctypes.c_int64),
                         _p(self.elem_tags, ctypes.c_int64))
        self.config_text = config_text
        self.node_bc = np.zeros(self.K * self.Np, np.uint8)
        if L.dgh_node_bc(self.h, config_text.encode(), _p(self.node_bc, ctypes.c_uint8)):
            raise ValueError(L.dgh_last_error().decode())
        ng = ctypes.c_int(0)
        tags = np.zeros(64, np.int32)
        refl = np.zeros(64, np.int32)
        sizes = np.zeros(64, np.int64)
        L.dgh_bc_groups(self.h, config_text.encode(), 64, ctypes.byref(ng), _p(tags, ctypes.c_int32),
                        _p(refl, ctypes.c_int32), _p(sizes, ctypes.c_int64))
        self.bc_groups = []
        for i in range(ng.value):
            nodes = np.zeros(int(sizes[i]), np.int64)
            L.dgh_group_nodes(self.h, int(tags[i]), _p(nodes, ctypes.c_int64))
            self.bc_groups.append((int(tags[i]), int(refl[i]), nodes))

    def write(self, path):
        if self.h is None:
            raise RuntimeError("the host mesh model was released")
        if host().dgh_write(self.h, path.encode()):
            raise RuntimeError(host().dgh_last_error().decode())

    def release(self):
        """Free the host-side mesh model (the element arrays stay); write() is then unavailable."""
        if getattr(self, "h", None) is not None:
            host().dgh_free(self.h)
            self.h = None

    def __del__(self):
        self.release()


def read_mesh(path, config_text):
    return Mesh(_check_handle(host().dgh_read(path.encode())), config_text)


def gen_rect(nx, ny, Lx, Ly, order, side_tags, names, config_text):
    st = np.array(side_tags, np.int32)
    return Mesh(_check_handle(host().dgh_gen_rect(nx, ny, Lx, Ly, order, _p(st, ctypes.c_int), names.encode())),
                config_text)


def gen_box(nx, ny, nz, Lx, Ly, Lz, order, jitter, seed, side_tags, names, config_text, block=None, block_tag=0):
    st = np.array(side_tags, np.int32)
    b0 = np.array(block[0] if block else (0, 0, 0), np.int32)
    b1 = np.array(block[1] if block else (0, 0, 0), np.int32)
    h = host().dgh_gen_box(nx, ny, nz, Lx, Ly, Lz, order, jitter, seed, _p(b0, ctypes.c_int), _p(b1, ctypes.c_int),
                           block_tag, _p(st, ctypes.c_int), names.encode())
    return Mesh(_check_handle(h), config_text)


def gen_line(n, L, order, side_tags, names, config_text):
    st = np.array(side_tags, np.int32)
    return Mesh(_check_handle(host().dgh_gen_line(n, L, order, _p(st, ctypes.c_int), names.encode())), config_text)


# ----------------------------------------------------------------------------- device library
FP64, FP32 = 0, 1
SIGN_REFERENCE, SIGN_STABLE = 0, 1
VARIANT_DEFAULT, VARIANT_LDS, VARIANT_MFMA = 0, 1, 2
# MFMA stage-kernel tile shapes (dgk_set_tile_shape): built-in, 4 waves x 16 elements, 8 waves x 32
# elements, row split (4 waves x 4 elements), 4 waves x 16 elements with the operator in global memory,
# row split over 8 waves x 8 elements, 4 waves x 16 elements compiled for 3, 4 or 5 waves per SIMD,
# the same for 4, 5 and 3 waves per SIMD with the epilogue operands loaded late
TILE_SHAPES = (0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11)
# fp32_flags: 0 = default FP32 form (compensated solution); PLAIN_STATE = plain FP32 solution;
# OPERATOR_CORRECTION = add the bf16 operator-residual MFMA
FP32_PLAIN_STATE, FP32_OPERATOR_CORRECTION = 1, 2
DGK_OK, DGK_ERR_INVALID, DGK_ERR_UNSUPPORTED, DGK_ERR_MESH, DGK_ERR_DEVICE, DGK_ERR_NONFINITE = range(6)


class _Mesh(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("order", ctypes.c_int), ("num_elements", ctypes.c_int64),
                ("node_xyz", P(ctypes.c_double)), ("node_tags", P(ctypes.c_int64)), ("node_bc", P(ctypes.c_uint8))]


class _Phys(ctypes.Structure):
    _fields_ = [("c0", ctypes.c_double), ("rho0", ctypes.c_double), ("v0", ctypes.c_double * 3)]


class _Opts(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("flux_sign", ctypes.c_int), ("device", ctypes.c_int),
                ("variant", ctypes.c_int), ("fp32_flags", ctypes.c_int), ("addr64", ctypes.c_int),
                ("reserved", ctypes.c_int * 4)]


class _Info(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("order", ctypes.c_int), ("Np", ctypes.c_int), ("Nf", ctypes.c_int),
                ("Nfp", ctypes.c_int), ("num_elements", ctypes.c_int64), ("num_faces", ctypes.c_int64),
                ("num_boundary_faces", ctypes.c_int64), ("precision", ctypes.c_int), ("variant", ctypes.c_int),
                ("fc", ctypes.c_double), ("device_bytes", ctypes.c_int64)]


DEV_SYMBOLS = ("dgk_create", "dgk_destroy", "dgk_last_error", "dgk_info_get", "dgk_set_state", "dgk_get_state",
               "dgk_set_sources", "dgk_rk4", "dgk_euler", "dgk_rhs", "dgk_sync", "dgk_check_finite",
               "dgk_connectivity", "dgk_rk4_timed", "dgk_variant_name", "dgk_batch_create", "dgk_batch_destroy",
               "dgk_batch_last_error", "dgk_batch_set_state", "dgk_batch_get_state", "dgk_batch_rk4",
               "dgk_batch_sync", "dgk_batch_groups", "dgk_batch_prepare", "dgk_batch_rk4_timed",
               "dgk_set_receivers", "dgk_get_receivers", "dgk_batch_set_receivers", "dgk_batch_get_receivers",
               "dgk_set_launch", "dgk_load_tuning", "dgk_batch_set_launch", "dgk_set_tile_shape",
               "dgk_batch_set_tile_shape", "dgk_stage_addressing")


def device():
    """Load libdgk.so. Raises when it is not built: the solver has no CPU fallback."""
    global _dev
    if _dev is None:
        if not os.path.exists(DEV_LIB):
            raise RuntimeError(f"{DEV_LIB} is not built (run __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(DEV_LIB)
        vp = ctypes.c_void_p
        L.dgk_create.argtypes = [P(_Mesh), P(_Phys), P(_Opts), P(vp)]
        L.dgk_destroy.argtypes = [vp]
        L.dgk_last_error.argtypes = [vp]
        L.dgk_last_error.restype = ctypes.c_char_p
        L.dgk_info_get.argtypes = [vp, P(_Info)]
        L.dgk_set_state.argtypes = [vp, P(ctypes.c_double)]
        L.dgk_get_state.argtypes = [vp, P(ctypes.c_double)]
        L.dgk_set_sources.argtypes = [vp, ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double)]
        L.dgk_rk4.argtypes = [vp, ctypes.c_int, ctypes.c_double, P(ctypes.c_double)]
        L.dgk_euler.argtypes = [vp, ctypes.c_int, ctypes.c_double, P(ctypes.c_double)]
        L.dgk_rhs.argtypes = [vp, P(ctypes.c_double), P(ctypes.c_double), ctypes.c_double]
        L.dgk_sync.argtypes = [vp]
        L.dgk_check_finite.argtypes = [vp]
        L.dgk_connectivity.argtypes = [vp] + [P(ctypes.c_int32)] * 4
        L.dgk_rk4_timed.argtypes = [vp, ctypes.c_int, ctypes.c_double, P(ctypes.c_double), P(ctypes.c_double)]
        L.dgk_variant_name.argtypes = [ctypes.c_int]
        L.dgk_variant_name.restype = ctypes.c_char_p
        L.dgk_batch_create.argtypes = [ctypes.c_int, P(_Mesh), P(_Phys), P(ctypes.c_double), P(_Opts), P(vp)]
        L.dgk_batch_destroy.argtypes = [vp]
        L.dgk_batch_last_error.argtypes = [vp]
        L.dgk_batch_last_error.restype = ctypes.c_char_p
        L.dgk_batch_set_state.argtypes = [vp, ctypes.c_int, P(ctypes.c_double)]
        L.dgk_batch_get_state.argtypes = [vp, ctypes.c_int, P(ctypes.c_double)]
        L.dgk_batch_rk4.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.dgk_batch_sync.argtypes = [vp]
        L.dgk_set_receivers.argtypes = [vp, ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double), ctypes.c_int]
        L.dgk_get_receivers.argtypes = [vp, P(ctypes.c_double), ctypes.c_int, P(ctypes.c_int)]
        L.dgk_batch_set_receivers.argtypes = [vp, ctypes.c_int, ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double),
                                              ctypes.c_int]
        L.dgk_batch_get_receivers.argtypes = [vp, ctypes.c_int, P(ctypes.c_double), ctypes.c_int, P(ctypes.c_int)]
        L.dgk_set_launch.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.dgk_load_tuning.argtypes = [ctypes.c_char_p]
        L.dgk_batch_set_launch.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.dgk_set_tile_shape.argtypes = [vp, ctypes.c_int]
        L.dgk_batch_set_tile_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.dgk_stage_addressing.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, P(ctypes.c_int)]
        if os.path.exists(TUNING_FILE) and not os.environ.get("DGK_NO_TUNING"):
            if L.dgk_load_tuning(TUNING_FILE.encode()):
                raise RuntimeError(f"tuning table {TUNING_FILE}: {L.dgk_last_error(None).decode()}")
        L.dgk_batch_prepare.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.dgk_batch_rk4_timed.argtypes = [vp, ctypes.c_int, P(ctypes.c_double)]
        L.dgk_batch_groups.argtypes = [vp, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int64),
                                       P(ctypes.c_int), P(ctypes.c_int64), P(ctypes.c_int64)]
        _dev = L
    return _dev


def stage_addressing(dim, order, precision, num_elements):
    """True when a context of this size addresses the stage kernel's arrays through 64-bit
    per-tile bases (some array reaches 4 GiB), False for 32-bit offsets (no device use)."""
    L = device()
    out = ctypes.c_int()
    st = L.dgk_stage_addressing(int(dim), int(order), int(precision), int(num_elements), ctypes.byref(out))
    if st:
        raise DGKError(st, L.dgk_last_error(None).decode())
    return bool(out.value)


class DGKError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"dgk status {status}: {msg}")
        self.status = status


class Solver:
    """One libdgk context: the reference's Mesh + solver state on the GPU."""

    def __init__(self, mesh, c0, rho0, v0=(0.0, 0.0, 0.0), precision=FP64, flux_sign=SIGN_REFERENCE, device_id=0,
                 variant=0, fp32_flags=0, addr64=False):
        L = device()
        self.mesh = mesh
        self.K, self.Np = mesh.K, mesh.Np
        self._xyz = np.ascontiguousarray(mesh.node_xyz, dtype=np.float64)
        self._tags = np.ascontiguousarray(mesh.node_tags, dtype=np.int64)
        self._bc = np.ascontiguousarray(mesh.node_bc, dtype=np.uint8)
        m = _Mesh(mesh.dim, mesh.order, mesh.K, _p(self._xyz), _p(self._tags, ctypes.c_int64),
                  _p(self._bc, ctypes.c_uint8))
        ph = _Phys(float(c0), float(rho0), (ctypes.c_double * 3)(*[float(x) for x in v0]))
        op = _Opts(int(precision), int(flux_sign), int(device_id), int(variant), int(fp32_flags), int(bool(addr64)))
        h = ctypes.c_void_p()
        st = L.dgk_create(ctypes.byref(m), ctypes.byref(ph), ctypes.byref(op), ctypes.byref(h))
        if st:
            raise DGKError(st, L.dgk_last_error(None).decode())
        self.h = h
        self.info = _Info()
        L.dgk_info_get(self.h, ctypes.byref(self.info))

    def _check(self, st):
        if st:
            raise DGKError(st, device().dgk_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            device().dgk_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_state(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.size == 4 * self.K * self.Np
        self._check(device().dgk_set_state(self.h, _p(u)))

    def get_state(self):
        u = np.zeros((4, self.K * self.Np))
        self._check(device().dgk_get_state(self.h, _p(u)))
        return u

    def set_sources(self, sources):
        src = np.ascontiguousarray(sources, dtype=np.float64).reshape(-1, 9)
        self._check(device().dgk_set_sources(self.h, len(src), _p(src), _p(self._xyz)))

    def rk4(self, nsteps, dt, t_start=None):
        ts = None if t_start is None else np.ascontiguousarray(t_start, dtype=np.float64)
        self._check(device().dgk_rk4(self.h, int(nsteps), float(dt), None if ts is None else _p(ts)))

    def euler(self, nsteps, dt, t_start=None):
        ts = None if t_start is None else np.ascontiguousarray(t_start, dtype=np.float64)
        self._check(device().dgk_euler(self.h, int(nsteps), float(dt), None if ts is None else _p(ts)))

    def rhs(self, u, dt):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros_like(u)
        self._check(device().dgk_rhs(self.h, _p(u), _p(out), float(dt)))
        return out

    def sync(self):
        self._check(device().dgk_sync(self.h))

    def check_finite(self):
        return device().dgk_check_finite(self.h) == 0

    def connectivity(self):
        n = self.K * self.info.Nf
        arrs = [np.zeros(n, np.int32) for _ in range(4)]
        self._check(device().dgk_connectivity(self.h, *[_p(a, ctypes.c_int32) for a in arrs]))
        return dict(zip(("nbr", "nface", "perm", "sign"), (a.reshape(self.K, -1) for a in arrs)))

    def set_receivers(self, xyz, capacity):
        """Record the pressure at points xyz (n x 3) after every step, up to `capacity` steps."""
        x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self._nrec = len(x)
        self._check(device().dgk_set_receivers(self.h, len(x), _p(x), _p(self._xyz), int(capacity)))

    def get_receivers(self):
        """Recorded pressure, shape (steps, n)."""
        L = device()
        rows = ctypes.c_int()
        self._check(L.dgk_get_receivers(self.h, None, 0, ctypes.byref(rows)))
        out = np.zeros((rows.value, getattr(self, "_nrec", 0)))
        self._check(L.dgk_get_receivers(self.h, _p(out), rows.value, ctypes.byref(rows)))
        return out

    def set_launch(self, grid_pct=100, tile_order=0):
        """Launch shape of the stage kernel (tuning only; results are unchanged)."""
        self._check(device().dgk_set_launch(self.h, int(grid_pct), int(tile_order)))

    def set_tile_shape(self, shape):
        """Tile shape of the MFMA stage kernel (TILE_SHAPES; 0 the built-in one). Raises
        DGKError (status DGK_ERR_UNSUPPORTED) when the context has no kernel of that shape."""
        self._check(device().dgk_set_tile_shape(self.h, int(shape)))

    def rk4_timed(self, nsteps, dt):
        s, t = ctypes.c_double(), ctypes.c_double()
        self._check(device().dgk_rk4_timed(self.h, int(nsteps), float(dt), ctypes.byref(s), ctypes.byref(t)))
        return s.value, t.value


SCHEDULE_PARALLEL, SCHEDULE_SEQUENTIAL = 0, 1


class Batch:
    """Many independent simulations advanced together (libdgk batch API).

    Simulations sharing (dim, order, physics) are fused into one device context; each keeps
    its own time step. States use the single-simulation layout u[4][K*Np].
    """

    def __init__(self, meshes, phys, dts, precision=FP64, flux_sign=SIGN_REFERENCE, device_id=0, variant=0,
                 fp32_flags=0):
        L = device()
        n = len(meshes)
        assert len(phys) == n and len(dts) == n
        self.meshes = list(meshes)
        self._keep = []
        marr = (_Mesh * n)()
        parr = (_Phys * n)()
        for i, (m, ph) in enumerate(zip(meshes, phys)):
            xyz = np.ascontiguousarray(m.node_xyz, dtype=np.float64)
            tags = np.ascontiguousarray(m.node_tags, dtype=np.int64)
            bc = np.ascontiguousarray(m.node_bc, dtype=np.uint8)
            self._keep += [xyz, tags, bc]
            marr[i] = _Mesh(m.dim, m.order, m.K, _p(xyz), _p(tags, ctypes.c_int64), _p(bc, ctypes.c_uint8))
            v0 = ph.get("v0", (0.0, 0.0, 0.0))
            parr[i] = _Phys(float(ph["c0"]), float(ph["rho0"]), (ctypes.c_double * 3)(*[float(x) for x in v0]))
        dta = np.ascontiguousarray(dts, dtype=np.float64)
        op = _Opts(int(precision), int(flux_sign), int(device_id), int(variant), int(fp32_flags))
        h = ctypes.c_void_p()
        st = L.dgk_batch_create(n, marr, parr, _p(dta), ctypes.byref(op), ctypes.byref(h))
        self._keep = None  # the library copies the meshes
        if st:
            raise DGKError(st, L.dgk_batch_last_error(None).decode())
        self.h = h

    def _check(self, st):
        if st:
            raise DGKError(st, device().dgk_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            device().dgk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_state(self, i, u):
        m = self.meshes[i]
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.size == 4 * m.K * m.Np
        self._check(device().dgk_batch_set_state(self.h, int(i), _p(u)))

    def get_state(self, i):
        m = self.meshes[i]
        u = np.zeros((4, m.K * m.Np))
        self._check(device().dgk_batch_get_state(self.h, int(i), _p(u)))
        return u

    def rk4(self, nsteps, schedule=SCHEDULE_PARALLEL):
        self._check(device().dgk_batch_rk4(self.h, int(nsteps), int(schedule)))

    def sync(self):
        self._check(device().dgk_batch_sync(self.h))

    def set_receivers(self, i, xyz, capacity):
        """Record member i's pressure at points xyz (n x 3) after every following step, for at least
        `capacity` steps; rows already recorded for other members of its group are kept."""
        m = self.meshes[i]
        x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        nx = np.ascontiguousarray(m.node_xyz, dtype=np.float64)
        if not hasattr(self, "_nrec"):
            self._nrec = {}
        self._nrec[i] = len(x)
        self._check(device().dgk_batch_set_receivers(self.h, int(i), len(x), _p(x), _p(nx), int(capacity)))

    def get_receivers(self, i):
        """Member i's recorded pressure, shape (steps, n)."""
        L = device()
        rows = ctypes.c_int()
        self._check(L.dgk_batch_get_receivers(self.h, int(i), None, 0, ctypes.byref(rows)))
        out = np.zeros((rows.value, getattr(self, "_nrec", {}).get(i, 0)))
        self._check(L.dgk_batch_get_receivers(self.h, int(i), _p(out), rows.value, ctypes.byref(rows)))
        return out

    def set_launch(self, group, grid_pct=100, tile_order=0):
        self._check(device().dgk_batch_set_launch(self.h, int(group), int(grid_pct), int(tile_order)))

    def set_tile_shape(self, group, shape):
        self._check(device().dgk_batch_set_tile_shape(self.h, int(group), int(shape)))

    def prepare(self, nsteps, schedule=SCHEDULE_PARALLEL):
        """Capture the graph for rk4(nsteps, schedule) ahead of time."""
        self._check(device().dgk_batch_prepare(self.h, int(nsteps), int(schedule)))

    def pick_schedule(self, steps=10, margin=0.97):
        """Schedule autotuning for this ensemble on this GPU: time `steps` captured steps with the
        groups as parallel graph branches and `steps` in sequence (the members advance 2*steps
        steps), and return (schedule, {schedule: seconds}). Parallel branches are taken only when
        clearly faster (below `margin` of the sequence), since their per-kernel timings overlap."""
        trial = {}
        for sched in (SCHEDULE_PARALLEL, SCHEDULE_SEQUENTIAL):
            self.prepare(steps, sched)
            self.sync()
            t0 = time.perf_counter()
            self.rk4(steps, sched)
            self.sync()
            trial[sched] = time.perf_counter() - t0
        best = SCHEDULE_PARALLEL if trial[SCHEDULE_PARALLEL] < margin * trial[SCHEDULE_SEQUENTIAL] else \
            SCHEDULE_SEQUENTIAL
        return best, trial

    def tune_groups(self, steps=3, grids=(100, 200, 400), orders=(0, 1), margin=0.98):
        """Tile-shape and launch-shape autotuning of every group at its own size on this GPU (the
        tuning table is measured on one mesh per (dim, order, precision); a group's size and mix of
        members are its own). Each candidate (tile shape, grid_pct, tile_order) is set on every
        group that has a kernel of that shape, and all groups are timed at once with event-timed
        stage launches (`steps` RK4 steps per candidate, which the members advance); each group
        keeps its fastest candidate, one of another tile shape than the built-in one only when it
        beats the built-in shape's best by `margin`. Results do not depend on the launch shape and
        agree with the built-in tile shape to rounding. Returns (choices, steps advanced): per
        group, (tile_shape, grid_pct, tile_order, stage_us) of the choice, which stays set."""
        n = len(self.groups())
        times = [dict() for _ in range(n)]
        advanced = 0
        for sh in TILE_SHAPES:
            ok = []
            for gi in range(n):
                try:
                    self.set_tile_shape(gi, sh)
                    ok.append(gi)
                except DGKError as e:
                    if e.status != DGK_ERR_UNSUPPORTED:
                        raise
            for g in grids if ok else ():
                for o in orders:
                    for gi in ok:
                        self.set_launch(gi, g, o)
                    ms = self.rk4_timed(steps)
                    advanced += steps
                    for gi in ok:
                        times[gi][(sh, g, o)] = ms[gi] * 1e3 / (4 * steps)
        out = []
        for gi in range(n):
            base = min((t, k) for k, t in times[gi].items() if k[0] == 0)
            best = min((t, k) for k, t in times[gi].items())
            t, (sh, g, o) = best if best[0] < margin * base[0] else base
            self.set_tile_shape(gi, sh)
            self.set_launch(gi, g, o)
            out.append((sh, g, o, t))
        return out, advanced

    def rk4_timed(self, nsteps):
        """Run nsteps per group with event-timed stage launches; summed stage ms per group."""
        n = len(self.groups())
        out = np.zeros(n)
        self._check(device().dgk_batch_rk4_timed(self.h, int(nsteps), _p(out)))
        return out

    def groups(self):
        L = device()
        g = ctypes.c_int()
        self._check(L.dgk_batch_groups(self.h, ctypes.byref(g), None, None, None, None, None, None))
        n = g.value
        d, o, s = (np.zeros(n, np.int32) for _ in range(3))
        e, f, fb = (np.zeros(n, np.int64) for _ in range(3))
        self._check(L.dgk_batch_groups(self.h, ctypes.byref(g), _p(d, ctypes.c_int), _p(o, ctypes.c_int),
                                       _p(e, ctypes.c_int64), _p(s, ctypes.c_int), _p(f, ctypes.c_int64),
                                       _p(fb, ctypes.c_int64)))
        return [dict(dim=int(d[i]), order=int(o[i]), elements=int(e[i]), sims=int(s[i]), faces=int(f[i]),
                     boundary_faces=int(fb[i])) for i in range(n)]
