"""GPU surrogate kernels against the CPU restatement (oracle/surrogate_ref.py).

The oracle rounds to bfloat16 exactly where the kernels do and accumulates in float64, so the
remaining differences are FP32 accumulation order and the occasional bf16 rounding flip of an
intermediate that the two accumulations put on different sides of a rounding boundary."""
import os
import sys

import numpy as np
import pytest

import surrogate

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import surrogate_ref as sr  # noqa: E402

pytestmark = pytest.mark.gpu

SMALL = dict(in_dim=6, n_freq=32, width=128, depth=3, out_dim=128, max_batch=128, freq_sigma=2.0)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def toy_data(n, out_dim, seed):
    """Smooth targets: a pulse whose delay is the source-receiver distance."""
    rng = np.random.default_rng(seed)
    x = rng.random((n, 6))
    d = np.linalg.norm(x[:, :3] - x[:, 3:], axis=1)
    s = np.linspace(0, 2, out_dim)
    t = np.exp(-((s[None, :] - d[:, None]) / 0.15) ** 2) / (0.3 + d[:, None])
    return x, t


def test_forward_matches_oracle():
    net = surrogate.Net(seed=3, **SMALL)
    p, B = net.get_params()
    p = p + np.float32(0.01) * np.random.default_rng(2).standard_normal(len(p)).astype(np.float32)  # non-zero biases
    net.set_params(p)
    x = np.random.default_rng(4).random((128, 6))
    y = net.forward(x)
    ref = sr.forward(p.astype(np.float64), B, x.astype(np.float32), net.shapes(), emulate=True)
    err = rel(y, ref)
    print(f"forward rel L2 vs oracle {err:.2e}")
    assert err <= 2e-3
    y1 = net.forward(x[:5])  # padded partial tile gives the same rows
    assert np.array_equal(y1, y[:5])


def test_gradients_match_oracle():
    net = surrogate.Net(seed=5, **SMALL)
    p, B = net.get_params()
    x, t = toy_data(128, SMALL["out_dim"], 6)
    w = 1.0 / (128 * (t * t).sum(axis=1))
    loss, g = net.gradients(x, t, w)
    rloss, rg = sr.loss_and_grad(p.astype(np.float64), B, x.astype(np.float32), t.astype(np.float32),
                                 w.astype(np.float32), net.shapes(), emulate=True)
    assert abs(loss - rloss) <= 1e-3 * rloss
    o = 0
    for l, (m, k) in enumerate(net.shapes()):
        for name, size in (("W", m * k), ("b", m)):
            e = rel(g[o:o + size], rg[o:o + size])
            print(f"layer {l} d{name}: rel L2 {e:.2e}")
            assert e <= 1e-2
            o += size
    assert rel(g, rg) <= 5e-3


def test_training_reduces_loss_and_generalises_on_smooth_targets():
    # Fourier scale 0.25: larger scales interpolate the 1792 training points but do not generalise
    # (CPU float64 training of the same network: held-out median 3% at 0.25, 15% at 0.5, 95% at 2)
    net = surrogate.Net(seed=7, **dict(SMALL, freq_sigma=0.25))
    x, t = toy_data(2048, SMALL["out_dim"], 8)
    w = 1.0 / (t * t).sum(axis=1)
    net.set_data(x[:1792], t[:1792], w[:1792])
    losses = net.train(3000, 128, 2e-3, seed=1)
    print(f"loss {losses[:10].mean():.3e} -> {losses[-50:].mean():.3e}")
    assert losses[-50:].mean() < 0.01 * losses[:10].mean()
    y = net.forward(x[1792:])
    err = surrogate.rel_l2(y, t[1792:])
    print(f"held-out median rel L2 {np.median(err):.3e}")
    assert np.median(err) < 0.08


def test_full_size_network_shape_and_query_time():
    net = surrogate.Net()  # 6 x 512 (5 hidden SiLU layers + output), 128 frequencies, 2048 samples
    assert net.nparam == 256 * 512 + 512 + 4 * (512 * 512 + 512) + 512 * 2048 + 2048
    ms = net.forward_ms(256, reps=20)
    print(f"full-size forward: {ms:.3f} ms per 256 queries")
    assert ms < 5.0


def test_room_surrogate_with_dg_fallback():
    """Learned part end to end on a small reverberant room (2 x 1.5 x 1 m, P3, 0.3 m pulse): DG
    impulse responses from the batched solver with receivers, two networks trained on the GPU to
    correct the closed-form image-source series of the box, a disagreement gate calibrated on
    held-out DG pairs and checked at the 2% target, and the DG fallback."""
    import time

    import dgk
    import workloads as wl

    L = (2.0, 1.5, 1.0)
    m, ph, dt, _ = wl.room_box(*L, 0.125, 3)
    sigma = 0.3
    rng = np.random.default_rng(1000)
    S, R, nsteps, stride = 64, 256, 2048, 4
    lo, hi = np.full(3, 0.2), np.array(L) - 0.2
    src = lo + (hi - lo) * rng.random((S, 3))
    rcv = lo + (hi - lo) * rng.random((S, R, 3))
    t0 = time.perf_counter()
    series = surrogate.dg_impulse_responses(m, ph, dt, src, rcv, nsteps, sigma, stride)
    t_dg = (time.perf_counter() - t0) / S  # one DG solve answers one source's queries
    # the batched data equal a single-context run of the same source
    s = dgk.Solver(m, ph["c0"], ph["rho0"], precision=dgk.FP32, flux_sign=dgk.SIGN_STABLE)
    u = np.zeros((4, m.K * m.Np))
    u[0] = np.exp(-((m.node_xyz - src[0]) ** 2).sum(axis=1) / (2 * sigma ** 2))
    s.set_state(u)
    s.set_receivers(rcv[0], nsteps)
    s.rk4(nsteps, dt)
    single = s.get_receivers()[stride - 1::stride].T
    assert rel(series[0], single) <= 1e-4
    X_src, X_rcv, T = np.repeat(src, R, axis=0), rcv.reshape(-1, 3), series.reshape(S * R, -1)
    perm = rng.permutation(S * R)
    te, ca, tr = perm[:512], perm[512:1024], perm[1024:]
    times = (np.arange(nsteps // stride) + 1) * stride * dt
    base = lambda s_, r_: surrogate.image_source_series(s_, r_, L, sigma, ph["c0"], times)  # noqa: E731
    sur = surrogate.Surrogate(np.zeros(3), np.array(L), out_dim=nsteps // stride, scale=float(np.abs(T).max()),
                              freq_sigma=0.1, base=base)
    losses = sur.fit(X_src[tr], X_rcv[tr], T[tr], 15000, batch=256, lr=2e-3)
    assert all(l[-50:].mean() < 0.1 * l[:10].mean() for l in losses)
    sur.calibrate(X_src[ca], X_rcv[ca], T[ca], quantile=0.9)
    y, est = sur.predict(X_src[te], X_rcv[te])
    err = surrogate.rel_l2(y, T[te])
    err_base = surrogate.rel_l2(base(X_src[te], X_rcv[te]), T[te])
    # the method without the network: answer with the nearest training pair (reciprocity-aware)
    Xtr = np.concatenate([X_src[tr], X_rcv[tr]], axis=1)
    nn_err = []
    for i in te:
        q = np.concatenate([X_src[i], X_rcv[i]])
        d = np.minimum(np.linalg.norm(Xtr - q, axis=1), np.linalg.norm(Xtr - np.concatenate([q[3:], q[:3]]), axis=1))
        nn_err.append(np.linalg.norm(T[tr][np.argmin(d)] - T[i]) / np.linalg.norm(T[i]))
    corr = float(np.corrcoef(est, err)[0, 1])
    t_q = sur.nets[0].forward_ms(256, 10) * 2 / 256 * 1e-3  # two networks per query
    tol = 0.02  # the stated target: answers within 2% relative L2 of the DG solve
    acc = est <= tol
    print(f"held-out rel L2: network median {np.median(err):.4f}, image-source base alone {np.median(err_base):.3f}, "
          f"nearest pair {np.median(nn_err):.3f}; below 2%: {(err <= tol).mean():.2f}; estimate vs true error corr "
          f"{corr:.2f}; gate at 2% accepts {acc.mean():.3f} (true error <= 2% in "
          f"{(err[acc] <= tol).mean() if acc.any() else float('nan'):.2f} of them); "
          f"query {t_q * 1e6:.2f} us vs DG {t_dg * 1e3:.1f} ms per source", flush=True)
    assert np.median(err) < 0.5 * np.median(nn_err)
    assert np.median(err) < 0.2 * np.median(err_base)  # what the network gains over the closed form
    assert corr > 0.3
    assert t_dg / t_q >= 1000
    assert acc.mean() >= 0.05 and (err[acc] <= tol).mean() >= 0.8  # the calibrated gate bounds what it accepts
    # fallback: declined queries are answered by DG exactly
    calls = []

    def dg_solve(sq, rq):
        i = int(np.argmin(np.linalg.norm(X_src[te] - sq, axis=1) + np.linalg.norm(X_rcv[te] - rq, axis=1)))
        calls.append(i)
        return T[te][i]

    svc = surrogate.QueryService(sur, dg_solve, tol=tol)
    ans, by_net = svc.query(X_src[te], X_rcv[te])
    assert len(calls) == int((~by_net).sum())
    assert np.array_equal(ans[~by_net], T[te][~by_net])
    assert np.array_equal(by_net, acc)


@pytest.mark.skip(reason="target not met: on the 20 x 15 x 8 m room with the block (order 3, 0.109 m pulse, 4096 "
                         "training pairs) the image-source base plus the networks reach a held-out median of 22% "
                         "relative L2, the base alone 22% (feature scales 0.1-3 and up to 20000 Adam steps: "
                         "0.220-0.231); the gate at 2% accepts no query and every query falls back to the DG "
                         "solve (dgfem-acoustic_amd/surrogate_room.py --room c5)")
def test_room_surrogate_config5_meets_two_percent():
    """The learned part's target on its own room: at a 2% tolerance the gate answers at least a
    fifth of the held-out queries, and at least 90% of those are within 2% of the DG solve."""
    import json
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, os.path.join(root, "dgfem-acoustic_amd", "surrogate_room.py"), "--room",
                          "c5"], stdout=subprocess.PIPE, text=True, check=True).stdout.strip().split("\n")[-1]
    g = json.loads(out)["gate"]
    assert g["accepted"] >= 0.2 and g["accepted_within_tol"] >= 0.9
