"""Source-to-receiver surrogate (ctypes binding of the dgs_* C ABI in libdgk.so) and the
DG-backed query service around it.

Network and kernels: include/dgk_surrogate.h, csrc/surrogate.hip. Training data are receiver
time series that the batched DG solver records (dgk.Batch with receivers); a query whose
error estimate -- the relative L2 disagreement of two independently trained networks --
exceeds the tolerance is answered by a DG solve instead.
"""
import ctypes

import numpy as np

import dgk

_P = ctypes.POINTER


class _Config(ctypes.Structure):
    _fields_ = [("in_dim", ctypes.c_int), ("n_freq", ctypes.c_int), ("width", ctypes.c_int), ("depth", ctypes.c_int),
                ("out_dim", ctypes.c_int), ("max_batch", ctypes.c_int), ("freq_sigma", ctypes.c_float),
                ("seed", ctypes.c_uint64)]


_bound = False


def _lib():
    global _bound
    L = dgk.device()
    if not _bound:
        vp = ctypes.c_void_p
        fp = _P(ctypes.c_float)
        L.dgs_create.argtypes = [_P(_Config), _P(vp)]
        L.dgs_destroy.argtypes = [vp]
        L.dgs_last_error.argtypes = [vp]
        L.dgs_last_error.restype = ctypes.c_char_p
        L.dgs_num_params.argtypes = [vp]
        L.dgs_num_params.restype = ctypes.c_int64
        L.dgs_get_params.argtypes = [vp, fp, fp]
        L.dgs_set_params.argtypes = [vp, fp, fp]
        L.dgs_forward.argtypes = [vp, ctypes.c_int, fp, fp]
        L.dgs_forward_timed.argtypes = [vp, ctypes.c_int, ctypes.c_int, _P(ctypes.c_double)]
        L.dgs_forward_noalloc.argtypes = [vp, ctypes.c_int, vp, vp]
        L.dgs_gradients.argtypes = [vp, ctypes.c_int, fp, fp, fp, fp, _P(ctypes.c_double)]
        L.dgs_set_data.argtypes = [vp, ctypes.c_int64, fp, fp, fp]
        L.dgs_train.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64, fp]
        _bound = True
    return L


SYMBOLS = ("dgs_create", "dgs_destroy", "dgs_last_error", "dgs_num_params", "dgs_get_params", "dgs_set_params",
           "dgs_forward", "dgs_forward_timed", "dgs_gradients", "dgs_set_data", "dgs_train", "dgs_forward_noalloc")


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_P(ctypes.c_float))


class Net:
    """One network on the device."""

    def __init__(self, in_dim=6, n_freq=128, width=512, depth=5, out_dim=2048, max_batch=256, freq_sigma=0.25,
                 seed=0):
        L = _lib()
        self.cfg = dict(in_dim=in_dim, n_freq=n_freq, width=width, depth=depth, out_dim=out_dim,
                        max_batch=max_batch, freq_sigma=freq_sigma, seed=seed)
        c = _Config(in_dim, n_freq, width, depth, out_dim, max_batch, freq_sigma, seed)
        h = ctypes.c_void_p()
        st = L.dgs_create(ctypes.byref(c), ctypes.byref(h))
        if st:
            raise RuntimeError(f"dgs_create: {L.dgs_last_error(None).decode()}")
        self.h = h
        self.nparam = L.dgs_num_params(h)

    def _check(self, st, what):
        if st:
            raise RuntimeError(f"{what}: {_lib().dgs_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            _lib().dgs_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def shapes(self):
        c = self.cfg
        dims = [2 * c["n_freq"]] + [c["width"]] * c["depth"] + [c["out_dim"]]
        return [(dims[i + 1], dims[i]) for i in range(c["depth"] + 1)]

    def get_params(self):
        p = np.zeros(self.nparam, np.float32)
        B = np.zeros((self.cfg["n_freq"], self.cfg["in_dim"]), np.float32)
        self._check(_lib().dgs_get_params(self.h, p.ctypes.data_as(_P(ctypes.c_float)),
                                          B.ctypes.data_as(_P(ctypes.c_float))), "dgs_get_params")
        return p, B

    def set_params(self, params=None, fourier=None):
        pa = _f(params) if params is not None else (None, None)
        fa = _f(fourier) if fourier is not None else (None, None)
        self._check(_lib().dgs_set_params(self.h, pa[1], fa[1]), "dgs_set_params")

    def forward(self, x):
        xa, xp = _f(x)
        y = np.zeros((len(xa), self.cfg["out_dim"]), np.float32)
        self._check(_lib().dgs_forward(self.h, len(xa), xp, y.ctypes.data_as(_P(ctypes.c_float))), "dgs_forward")
        return y

    def forward_device(self, rows, x_ptr, y_ptr):
        """Forward pass on device buffers (addresses x_ptr [rows][in_dim], y_ptr [rows][out_dim], FP32)
        without any device allocation (dgs_forward_noalloc)."""
        self._check(_lib().dgs_forward_noalloc(self.h, int(rows), ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr)),
                    "dgs_forward_noalloc")

    def forward_ms(self, rows, reps=20):
        ms = ctypes.c_double()
        self._check(_lib().dgs_forward_timed(self.h, int(rows), int(reps), ctypes.byref(ms)), "dgs_forward_timed")
        return ms.value

    def gradients(self, x, t, w):
        xa, xp = _f(x)
        ta, tp = _f(t)
        wa, wp = _f(w)
        g = np.zeros(self.nparam, np.float32)
        loss = ctypes.c_double()
        self._check(_lib().dgs_gradients(self.h, len(xa), xp, tp, wp, g.ctypes.data_as(_P(ctypes.c_float)),
                                         ctypes.byref(loss)), "dgs_gradients")
        return loss.value, g

    def set_data(self, x, t, w):
        xa, xp = _f(x)
        ta, tp = _f(t)
        wa, wp = _f(w)
        self._check(_lib().dgs_set_data(self.h, len(xa), xp, tp, wp), "dgs_set_data")

    def train(self, steps, batch, lr, seed=0):
        losses = np.zeros(steps, np.float32)
        self._check(_lib().dgs_train(self.h, int(steps), int(batch), float(lr), int(seed),
                                     losses.ctypes.data_as(_P(ctypes.c_float))), "dgs_train")
        return losses


def dg_impulse_responses(mesh, phys, dt, sources, receivers, nsteps, sigma, stride=1, precision=dgk.FP32,
                         flux_sign=dgk.SIGN_STABLE, per_batch=32):
    """Receiver pressure series of Gaussian pulses (width sigma) released at each source.

    sources: (S, 3); receivers: (S, R, 3) -- receiver set of each source. One batch member per
    source (the same mesh repeated), `per_batch` members per batch. Returns (S, R, nsteps // stride)
    sampled after steps stride, 2 stride, ...
    """
    sources = np.asarray(sources, np.float64)
    receivers = np.asarray(receivers, np.float64)
    S, R = receivers.shape[:2]
    out = np.zeros((S, R, nsteps // stride))
    x = mesh.node_xyz
    for s0 in range(0, S, per_batch):
        idx = list(range(s0, min(S, s0 + per_batch)))
        b = dgk.Batch([mesh] * len(idx), [phys] * len(idx), [dt] * len(idx), precision=precision,
                      flux_sign=flux_sign)
        for j, i in enumerate(idx):
            u = np.zeros((4, mesh.K * mesh.Np))
            u[0] = np.exp(-((x - sources[i]) ** 2).sum(axis=1) / (2 * sigma ** 2))
            b.set_state(j, u)
            b.set_receivers(j, receivers[i], nsteps)
        b.rk4(nsteps)
        for j, i in enumerate(idx):
            rec = b.get_receivers(j)
            out[i] = rec[stride - 1::stride][:nsteps // stride].T
        b.close()
    return out


def gaussian_pulse_free_field(r, t, sigma, c0):
    """Pressure at distance r and time t of a Gaussian pressure pulse exp(-|x|^2/(2 sigma^2)) released
    at rest in free space (linear acoustics, v0 = 0): the spherical d'Alembert solution
    p = [(r - c t) g(r - c t) + (r + c t) g(r + c t)] / (2 r) with g(s) = exp(-s^2/(2 sigma^2))."""
    a, b = r - c0 * t, r + c0 * t
    return (a * np.exp(-a * a / (2 * sigma ** 2)) + b * np.exp(-b * b / (2 * sigma ** 2))) / (2 * r)


def image_source_series(src, rcv, L, sigma, c0, t, rigid=(True, True, True, True, True, False)):
    """Receiver pressure series of the Gaussian pulse in the box [0,Lx]x[0,Ly]x[0,Lz] by images:
    walls marked rigid (x=0, x=Lx, y=0, y=Ly, z=0, z=Lz) mirror the source with the same sign, the
    others are taken as perfectly absorbing (no image). Obstacles inside the box are ignored.
    src, rcv: (P, 3); t: (T,) times. Returns (P, T)."""
    src = np.atleast_2d(np.asarray(src, np.float64))
    rcv = np.atleast_2d(np.asarray(rcv, np.float64))
    t = np.asarray(t, np.float64)
    reach = c0 * t.max() + 8 * sigma
    out = np.zeros((len(src), len(t)))

    def axis_images(s, Lk, lo_rigid, hi_rigid):
        """image coordinates along one axis (within `reach` of the box)"""
        if lo_rigid and hi_rigid:  # period 2L lattice: 2nL + s and 2nL - s
            n = int(np.ceil(reach / (2 * Lk))) + 1
            k = np.arange(-n, n + 1)
            return np.concatenate([2 * k * Lk + s, 2 * k * Lk - s])
        pts = [s]
        if lo_rigid:
            pts.append(-s)
        if hi_rigid:
            pts.append(2 * Lk - s)
        return np.array(pts)

    for i in range(len(src)):
        ax = [axis_images(src[i, k], L[k], rigid[2 * k], rigid[2 * k + 1]) for k in range(3)]
        d = [rcv[i, k] - ax[k] for k in range(3)]
        d = [dk[np.abs(dk) <= reach] for dk in d]
        r = np.sqrt(d[0][:, None, None] ** 2 + d[1][None, :, None] ** 2 + d[2][None, None, :] ** 2).ravel()
        r = np.maximum(r[r <= reach], 1e-9)
        out[i] = gaussian_pulse_free_field(r[:, None], t[None, :], sigma, c0).sum(axis=0)
    return out


def rel_l2(a, b):
    """Per-row relative L2 error |a - b| / |b|."""
    return np.linalg.norm(a - b, axis=-1) / np.maximum(np.linalg.norm(b, axis=-1), 1e-30)


class Surrogate:
    """Two networks trained on the same data with different seeds. The answer is their mean; the
    error estimate is their relative disagreement. With a `base` model (a callable (src, rcv) ->
    series, e.g. image_source_series), the networks learn the correction to it: answer = base +
    mean network output."""

    def __init__(self, lo, hi, out_dim, scale, seeds=(1, 2), base=None, **net_kw):
        self.lo = np.asarray(lo, np.float64)
        self.hi = np.asarray(hi, np.float64)
        self.scale = float(scale)
        self.base = base
        self.cal = None  # scale of the disagreement that bounds the true error (calibrate())
        self.nets = [Net(out_dim=out_dim, seed=s, **net_kw) for s in seeds]

    def _base(self, src, rcv, n):
        return self.base(src, rcv) if self.base is not None else np.zeros((len(np.atleast_2d(src)), n))

    def encode(self, src, rcv):
        """Coordinates normalised to [0,1] by the domain box."""
        src = np.atleast_2d(src)
        rcv = np.atleast_2d(rcv)
        return np.concatenate([(src - self.lo) / (self.hi - self.lo), (rcv - self.lo) / (self.hi - self.lo)], axis=1)

    def fit(self, src, rcv, series, steps, batch=256, lr=1e-3, seed=0, decay=(1.0, 0.3, 0.1)):
        """Train every network on (src, rcv) -> series; the learning rate steps down through
        `decay` over equal thirds of the run."""
        x = self.encode(src, rcv)
        full = np.asarray(series, np.float64) / self.scale
        t = full - self._base(src, rcv, full.shape[1]) / self.scale
        # relative L2 loss per sample (relative to the full series); pairs the pulse has not reached
        # within the window carry almost no signal, so their weight is capped at that of a series
        # 1% of the largest norm
        nrm2 = (full * full).sum(axis=1)
        w = 1.0 / np.maximum(nrm2, 1e-4 * nrm2.max())
        out = []
        for k, n in enumerate(self.nets):
            n.set_data(x, t, w)
            parts = [steps // len(decay)] * len(decay)
            parts[-1] += steps - sum(parts)
            out.append(np.concatenate([n.train(ns, batch, lr * f, seed + 101 * k + 7 * j)
                                       for j, (ns, f) in enumerate(zip(parts, decay)) if ns > 0]))
        return out

    def predict(self, src, rcv):
        """(mean prediction, error estimate per query). With two networks whose errors are
        independent and of equal size e, their difference has size sqrt(2) e and their mean's
        error is e / sqrt(2): the estimate of the mean's relative error is half the relative
        disagreement."""
        x = self.encode(src, rcv)
        ys = [n.forward(x).astype(np.float64) * self.scale for n in self.nets]
        if self.base is not None:
            b = self._base(src, rcv, ys[0].shape[1])
            ys = [y + b for y in ys]
        mean = sum(ys) / len(ys)
        dis = rel_l2(ys[0], ys[1]) if len(ys) > 1 else np.zeros(len(mean))
        return mean, dis * (self.cal if self.cal is not None else 0.5)

    def calibrate(self, src, rcv, series, quantile=0.9):
        """Scale the disagreement so that on these held-out pairs (DG truths not used in
        training) the estimate bounds the true relative error in `quantile` of the cases."""
        self.cal = None
        y, est = self.predict(src, rcv)
        dis = 2.0 * est
        err = rel_l2(y, np.asarray(series, np.float64))
        self.cal = float(np.quantile(err / np.maximum(dis, 1e-30), quantile))
        return self.cal


class QueryService:
    """Receiver queries answered by the surrogate when its error estimate is within `tol`,
    otherwise by a DG solve (`dg_solve(src, rcv) -> series`)."""

    def __init__(self, surrogate, dg_solve, tol=0.02):
        self.s = surrogate
        self.dg_solve = dg_solve
        self.tol = tol

    def query(self, src, rcv):
        y, est = self.s.predict(src, rcv)
        use_dg = est > self.tol
        for i in np.nonzero(use_dg)[0]:
            y[i] = self.dg_solve(np.atleast_2d(src)[i], np.atleast_2d(rcv)[i])
        return y, ~use_dg
