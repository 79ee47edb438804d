"""Learned receiver surrogate for the shoebox room at order 3, end to end on one GPU.

Room: 20 x 15 x 8 m, a 2 x 2 x 1 m rigid block on the floor at a seeded grid-aligned spot
(seed 4), Kuhn tets of 0.25 m (920,064 elements), order 3, walls/floor/block reflecting, ceiling
absorbing, c0 = 343. Impulse responses: Gaussian pulses of the 500 Hz-equivalent width
sigma = c0/(2 pi 500) released at seeded source points; pressure recorded at seeded receiver
points after every one of 2048 RK4 steps (FP32 batched DG solver, one batch member per source).
Data: 40 sources x 128 receivers = 5120 pairs (seed 1000): 4096 training, 512 calibration,
512 held out.
Network: 6 x 512 (5 SiLU layers of 512 + a 2048-sample output layer), 128 Fourier frequencies,
BF16 MFMA kernels with FP32 accumulation; two networks (seeds 1, 2) whose disagreement is the
error estimate; queries above the 2% estimate fall back to a DG solve. By default the networks
learn the correction to the closed-form image-source series of the box (surrogate.py:
image_source_series; rigid walls and floor mirrored, the absorbing ceiling and the block left
out), and the JSON line states that base's own held-out error beside the surrogate's.

Prints one JSON line with the held-out error distribution, the gate's acceptance, and the query
time against the DG solve time.

A second configuration (--room small) runs the same pipeline where the mapping is learnable
from a few thousand DG runs' worth of pairs: a 2 x 1.5 x 1 m room at order 3 (9,216 tets) with
a 0.3 m pulse, 64 sources x 256 receivers, 512 samples (every 4th of 2048 steps). Both split the
pairs into training, 512 calibration pairs (which scale the gate's estimate so that it bounds the
true error in 90% of them) and 512 held-out pairs.

Usage: python dgfem-acoustic_amd/surrogate_room.py [--room c5|small] [--out PATH]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import dgk  # noqa: E402
import surrogate  # noqa: E402
import workloads as wl  # noqa: E402


def log(msg):
    print(f"[surrogate_room] {msg}", file=sys.stderr, flush=True)


def points(rng, n, sigma, block_box, L):
    """Uniform points at least 3 sigma from the walls and outside the block (grown by 3 sigma)."""
    lo, hi = np.full(3, 3 * sigma), np.asarray(L) - 3 * sigma
    out = []
    while len(out) < n:
        p = lo + (hi - lo) * rng.random(3)
        if np.all(p > block_box[0] - 3 * sigma) and np.all(p < block_box[1] + 3 * sigma):
            continue
        out.append(p)
    return np.array(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--room", choices=("c5", "small"), default="c5",
                    help="c5: the 20 x 15 x 8 m room with a block at order 3, 500 Hz pulse width; small: a "
                         "2 x 1.5 x 1 m room at order 3 with a 0.3 m (180 Hz-equivalent) pulse")
    ap.add_argument("--sources", type=int, default=0)
    ap.add_argument("--receivers", type=int, default=0)
    ap.add_argument("--nsteps", type=int, default=2048)
    ap.add_argument("--stride", type=int, default=0, help="keep every stride-th sample")
    ap.add_argument("--steps", default="0", help="training steps per network (a comma-separated list: each in turn)")
    ap.add_argument("--lr", type=float, default=0.0, help="learning rate (default 2e-3 small, 1e-3 c5)")
    ap.add_argument("--freq-sigma", default="0.1",
                    help="scale of the Fourier feature frequencies (per unit of the room-normalised "
                         "coordinates); a comma-separated list trains and reports each in turn on the same data")
    ap.add_argument("--tol", type=float, default=0.02)
    ap.add_argument("--out", default="")
    ap.add_argument("--dump", default="", help="write sources, receivers and DG series (float32) to this .npz")
    ap.add_argument("--base", choices=("image", "none"), default="image",
                    help="image: the networks learn the correction to the image-source series of the box (rigid "
                         "walls and floor, absorbing ceiling, no obstacle); none: the series themselves")
    args = ap.parse_args()
    small = args.room == "small"
    S = args.sources or (64 if small else 40)
    R = args.receivers or (256 if small else 128)
    stride = args.stride or (4 if small else 1)
    step_list = [int(x) for x in str(args.steps).split(",")]
    step_list = [x or (15000 if small else 6000) for x in step_list]
    lr = args.lr or (2e-3 if small else 1e-3)
    t0 = time.perf_counter()
    if small:
        L = (2.0, 1.5, 1.0)
        m, ph, dt, _ = wl.room_box(*L, 0.125, 3)
        sigma = 0.3
        block_box = (np.full(3, -1.0), np.full(3, -1.0))  # no obstacle
        margin = 0.2
    else:
        L = (20.0, 15.0, 8.0)
        h = 0.25
        nx, ny = int(round(L[0] / h)), int(round(L[1] / h))
        brng = np.random.Generator(np.random.MT19937(4))
        bx, by = int(2 / h), int(2 / h)
        i0, j0 = int(brng.integers(0, nx - bx)), int(brng.integers(0, ny - by))
        block = ((i0, j0, 0), (i0 + bx, j0 + by, int(round(1 / h))))
        block_box = (np.array([i0 * h, j0 * h, 0.0]), np.array([(i0 + bx) * h, (j0 + by) * h, 1.0]))
        m, ph, dt, sigma = wl.room_box(*L, h, 3, block=block)
        margin = None
    m.release()
    log(f"room K={m.K} Np={m.Np} dt={dt:.3e} sigma={sigma:.4f} ({time.perf_counter() - t0:.1f} s)")
    rng = np.random.default_rng(1000)
    src = points(rng, S, sigma if margin is None else margin / 3, block_box, L)
    rcv = np.stack([points(rng, R, sigma if margin is None else margin / 3, block_box, L) for _ in range(S)])
    t0 = time.perf_counter()
    series = surrogate.dg_impulse_responses(m, ph, dt, src, rcv, args.nsteps, sigma, stride, per_batch=12)
    t_data = time.perf_counter() - t0
    log(f"DG data: {S} sources x {R} receivers x {args.nsteps // stride} samples in {t_data:.1f} s")
    if args.dump:
        np.savez_compressed(args.dump, src=src, rcv=rcv, series=series.astype(np.float32), dt=dt, sigma=sigma,
                            stride=stride, L=np.array(L), block_lo=block_box[0], block_hi=block_box[1])
    # one plain DG solve (single context, one source, all its receivers) for the per-query cost
    s = dgk.Solver(m, ph["c0"], ph["rho0"], precision=dgk.FP32, flux_sign=dgk.SIGN_STABLE)
    u = np.zeros((4, m.K * m.Np))
    u[0] = np.exp(-((m.node_xyz - src[0]) ** 2).sum(axis=1) / (2 * sigma ** 2))
    s.set_state(u)
    s.set_receivers(rcv[0], args.nsteps)
    s.sync()
    t0 = time.perf_counter()
    s.rk4(args.nsteps, dt)
    s.sync()
    t_dg = time.perf_counter() - t0
    single = s.get_receivers()[stride - 1::stride].T
    s.close()
    consistency = float(np.linalg.norm(single - series[0]) / np.linalg.norm(single))
    X_src, X_rcv, T = np.repeat(src, R, axis=0), rcv.reshape(-1, 3), series.reshape(S * R, -1)
    perm = rng.permutation(S * R)
    te, ca, tr = perm[:512], perm[512:1024], perm[1024:]
    times = (np.arange(args.nsteps // stride) + 1) * stride * dt
    base = None
    if args.base == "image":
        base = lambda s_, r_: surrogate.image_source_series(s_, r_, L, sigma, ph["c0"], times)  # noqa: E731
    if args.out:
        open(args.out, "w").close()
    te0 = te
    for freq_sigma, steps in [(float(f), n) for f in str(args.freq_sigma).split(",") for n in step_list]:
        te = te0
        sur = surrogate.Surrogate(np.zeros(3), np.array(L), out_dim=args.nsteps // stride, scale=float(np.abs(T).max()),
                                  freq_sigma=freq_sigma, base=base)
        t0 = time.perf_counter()
        losses = sur.fit(X_src[tr], X_rcv[tr], T[tr], steps, batch=256, lr=lr)
        t_train = time.perf_counter() - t0
        # relative errors (held out, calibration, training) are taken over the pairs the pulse reaches
        # within the window (series norm above 1e-3 of the largest); the silent rest have no relative
        # error to speak of and are counted separately
        loud = np.linalg.norm(T, axis=1) > 1e-3 * np.linalg.norm(T, axis=1).max()
        audible = loud[te]
        te, ca_loud, tr_loud = te[audible], ca[loud[ca]], tr[loud[tr]][:512]
        cal = sur.calibrate(X_src[ca_loud], X_rcv[ca_loud], T[ca_loud], quantile=0.9)
        y, est = sur.predict(X_src[te], X_rcv[te])
        err = surrogate.rel_l2(y, T[te])
        err_base = surrogate.rel_l2(base(X_src[te], X_rcv[te]), T[te]) if base is not None else None
        ytr, _ = sur.predict(X_src[tr_loud], X_rcv[tr_loud])
        err_tr = surrogate.rel_l2(ytr, T[tr_loud])
        t_q = sur.nets[0].forward_ms(256, 20) * len(sur.nets) / 256 * 1e-3
        acc = est <= args.tol
        out = {
            "room": {"name": args.room, "elements": m.K, "order": 3, "dofs": 4 * m.K * m.Np, "dt": dt,
                     "pulse_sigma": sigma, "samples": args.nsteps // stride, "sample_stride": stride},
            "data": {"sources": S, "receivers_per_source": R, "pairs_train": len(tr), "pairs_calibration": len(ca),
                     "pairs_held_out": 512, "held_out_audible": int(audible.sum()),
                     "calibration_audible": len(ca_loud), "dg_seconds": t_data,
                     "batched_vs_single_rel_l2": consistency},
            "network": {"params": sur.nets[0].nparam, "freq_sigma": freq_sigma, "train_steps": steps,
                        "lr": lr, "train_seconds": t_train, "final_loss": [float(l[-50:].mean()) for l in losses]},
            "held_out_rel_l2": {"median": float(np.median(err)), "mean": float(err.mean()),
                                "p10": float(np.percentile(err, 10)), "frac_below_tol": float((err <= args.tol).mean())},
            "base": args.base,
            "held_out_rel_l2_base_alone": None if err_base is None else
            {"median": float(np.median(err_base)), "p10": float(np.percentile(err_base, 10)),
             "frac_below_tol": float((err_base <= args.tol).mean())},
            "train_rel_l2_median": float(np.median(err_tr)),
            "gate": {"tol": args.tol, "calibration_scale": cal, "accepted": float(acc.mean()),
                     "accepted_err_median": float(np.median(err[acc])) if acc.any() else None,
                     "accepted_within_tol": float((err[acc] <= args.tol).mean()) if acc.any() else None,
                     "accepted_err_max": float(err[acc].max()) if acc.any() else None,
                     "corr_estimate_error": float(np.corrcoef(est, err)[0, 1])},
            "gate_by_tolerance": {f"{tol:g}": {"accepted": float((est <= tol).mean()),
                                               "accepted_within_tol": float((err[est <= tol] <= tol).mean())
                                               if (est <= tol).any() else None}
                                  for tol in (0.02, 0.03, 0.05, 0.1)},
            "timing": {"query_us": t_q * 1e6, "dg_solve_s": t_dg, "speedup": t_dg / t_q},
        }
        line = json.dumps(out)
        print(line, flush=True)
        if args.out:
            with open(args.out, "a") as f:
                f.write(line + "\n")


if __name__ == "__main__":
    main()
