"""Benchmark: DOF-RK-stage updates/s of the DG hot path on one MI355X.

Workload: the 3D order-3 drop-in case -- unit cube, 32^3 cells Kuhn-split into 196,608
tetrahedra, interior nodes jittered by 10% of h (seed 2), 4 fields (p, vx, vy, vz), mean
flow v0 = (50, 0, 0) m/s, absorbing boundaries, Gaussian pulse (sigma 0.04) at a seeded
interior point, dt = 0.1 h_min / ((c0 + |v0|)(2p+1)) (CFL 0.1 on the DG node spacing). Synthetic data: the mesh and state are
generated here; inputs are resident in HBM before the timed region.

One "step" is one classic RK4 step (4 fused stage launches). The FP64 drop-in path and
the FP32 path run the same number of steps from the same state; the FP32 relative L2
error per field against the FP64 result (which matches the CPU oracle to <= 1e-12) is
reported with the rates.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dgfem-acoustic_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import dgk  # noqa: E402
import workloads as wl  # noqa: E402

METRIC = "DOF-RK-stage updates/s: FP64 plain vs FP32-MFMA batched at stated L2 err"
HBM_PEAK = 8.0e12
FP64_PEAK = 78.6e12
FP32_PEAK = 157.3e12


def algorithmic(m, info, word):
    """Algorithmic bytes and flops per DOF-stage of the fused stage (4 fields, affine elements).

    bytes: per RK stage the state moves on average 4 words per DOF (stage 1: read u, write acc and
    the next stage input; stages 2-3: read stage input, u and acc, write acc and next input;
    stage 4: read stage input and acc, write u), plus the neighbour trace of every interior face
    read by each side (2 F_int Nfp / (K Np) words per DOF), plus per element the inverse Jacobian
    (9 words) and per element face an int32 neighbour and an int32 permutation code.
    flops: volume d*Np*Np + lift Nf*Nfp*Np multiply-adds per field and element, the contravariant
    flux, the nodal face flux and the RK update.
    """
    d, Np, Nf, Nfp, K = m.dim, m.Np, info.Nf, info.Nfp, m.K
    nEq = 4
    F_int = info.num_faces - info.num_boundary_faces
    words = 4 + 2 * F_int * Nfp / (K * Np)
    geo = (9 * word + 8 * Nf) / (nEq * Np)
    bytes_per = word * words + geo
    flops_per = 2 * d * Np + 2 * Nf * Nfp + (2 * d * 3) + (2 * F_int + info.num_boundary_faces) * Nfp * 16 / (K * Np * nEq) + 4
    return bytes_per, flops_per


def run_form(m, ph, dt, u0, precision, steps, warmup, variant=0, fp32_flags=0):
    s = dgk.Solver(m, ph["c0"], ph["rho0"], v0=ph["v0"], precision=precision, flux_sign=dgk.SIGN_STABLE,
                   variant=variant, fp32_flags=fp32_flags)
    s.set_state(u0)
    s.rk4(warmup, dt)
    s.sync()
    t0 = time.perf_counter()
    s.rk4(steps, dt)
    s.sync()
    wall = time.perf_counter() - t0
    # per-launch kernel time with HIP events on the context's stream (same work, same state path)
    nk = max(10, min(steps, 200))
    stage_ms, total_ms = s.rk4_timed(nk, dt)
    state = s.get_state()
    finite = bool(np.isfinite(state).all())
    return s, wall, stage_ms / (4 * nk), state, finite, steps + warmup + nk


def cpu_baseline(m, ph, dt, u0):
    import oracle_binding as ob

    threads = os.cpu_count() or 1
    o = ob.Oracle(m, ph["c0"], ph["rho0"], v0=ph["v0"], sign_mode=ob.SIGN_STABLE, nthreads=threads,
                  factorized_only=True)
    t0 = time.perf_counter()
    o.step(u0, 1, dt)
    sec = time.perf_counter() - t0
    dof = 4 * m.K * m.Np
    return {"value": dof * 4 / sec, "unit": "DOF-RK-stage updates/s", "cores": threads, "kind": "port",
            "sample": f"1 RK4 step of the full 196,608-tet workload with the factorized OpenMP oracle ({sec:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp64,fp32", help="comma list of fp64,fp32")
    ap.add_argument("--variant", type=int, default=0, help="stage-kernel variant (0 = default)")
    ap.add_argument("--fp32-flags", type=int, default=0, help="FP32 form switches (dgk.FP32_* bits, 0 = full form)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="first run a plain device copy of MB MiB (reference dispatch for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the sharding design is in DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy calibration: {args.copy_calibration} MiB read + {args.copy_calibration} MiB written",
              file=sys.stderr, flush=True)
        del src, dst
    m, ph, dt, u0 = wl.cube_3d()
    dof = 4 * m.K * m.Np
    res = {}
    forms = [f for f in (("fp64", dgk.FP64), ("fp32", dgk.FP32)) if f[0] in args.forms.split(",")]
    for name, prec in forms:
        print(f"[bench] {name}: {args.warmup} warmup + {args.steps} timed RK4 steps, K={m.K}", file=sys.stderr,
              flush=True)
        s, wall, kern_ms, state, finite, nsteps_total = run_form(m, ph, dt, u0, prec, args.steps, args.warmup,
                                                                 args.variant, args.fp32_flags)
        info = s.info
        rate = dof * 4 * args.steps / wall
        word = 8 if prec == dgk.FP64 else 4
        bpd, fpd = algorithmic(m, info, word)
        kern_rate = dof / (kern_ms * 1e-3)
        ach_bw = bpd * dof / (kern_ms * 1e-3)
        res[name] = dict(rate=rate, wall=wall, kern_ms=kern_ms, state=state, finite=finite, bpd=bpd, fpd=fpd,
                         ach_bw=ach_bw, kern_rate=kern_rate, variant=dgk.device().dgk_variant_name(info.variant).decode(),
                         steps_total=nsteps_total)
        s.close()
        print(f"[bench] {name}: {rate:.4g} DOF-stage/s wall, stage kernel {kern_ms * 1e3:.1f} us "
              f"({ach_bw / 1e9:.0f} GB/s algorithmic)", file=sys.stderr, flush=True)
    if len(res) < 2:
        for k, r in res.items():
            print(json.dumps({"form": k, "rate": r["rate"], "kernel_avg_us": r["kern_ms"] * 1e3}))
        return
    g64, g32 = res["fp64"]["state"], res["fp32"]["state"]
    names = ["p", "vx", "vy", "vz"]
    errs = {names[q]: float(np.linalg.norm(g32[q] - g64[q]) / max(np.linalg.norm(g64[q]), 1e-300)) for q in range(4)}
    r32, r64 = res["fp32"], res["fp64"]

    def roof(r):
        return {"bound": "hbm", "achieved": round(r["ach_bw"] / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(r["ach_bw"] / HBM_PEAK, 4), "traffic": None,
                "kernel": ("k_stage_lds" if r["variant"] == "lds-tile" else "k_stage_mfma") + f" ({r['variant']})",
                "kernel_avg_us": round(r["kern_ms"] * 1e3, 2),
                "algorithmic_bytes_per_dof_stage": round(r["bpd"], 2),
                "flop_per_dof_stage": round(r["fpd"], 1)}

    out = {
        "metric": METRIC,
        "value": r32["rate"],
        "unit": "DOF-RK-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["wall"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "cube3d-p3: unit cube, 32^3 Kuhn cells (196,608 tets), 10% jitter, order 3, "
                               "4 fields, v0=(50,0,0), absorbing walls, Gaussian pulse",
                   "elements": m.K, "order": 3, "dofs": dof, "dt": dt, "flux_sign": "stable",
                   "form": "fp32 MFMA arithmetic, fp32 solution with bf16 low word (compensated), fp64 geometry "
                           "rounded once", "fp32_flags": args.fp32_flags},
        "fp32_rel_l2_vs_fp64": {"after_steps": r32["steps_total"], **errs},
        "fp64_plain": {"value": r64["rate"], "ms_per_step": r64["wall"] * 1e3 / args.steps, "dtype": "fp64",
                       "roofline": roof(r64), "finite": r64["finite"]},
        "fp32_over_fp64": r32["rate"] / r64["rate"],
        "roofline": roof(r32),
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        print("[bench] cpu baseline (oracle, factorized, all host threads)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(m, ph, dt, u0)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
