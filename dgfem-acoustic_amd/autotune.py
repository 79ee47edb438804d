"""Autotuner for the stage kernel's tile shape and launch shape (the batch schedule is picked
per ensemble by dgk.Batch.pick_schedule).

For each (dim, order, precision) it builds a representative mesh (about `dofs` unknowns: a
structured 2D rectangle or a 3D Kuhn box), times every candidate with HIP events around each
stage launch (dgk_rk4_timed) and keeps the fastest (a tile shape other than the built-in one only
when it beats the built-in shape's best by 2%):

  variant     thread-per-node kernel (1, FP64 only) or MFMA kernel (2)
  tile_shape  the MFMA kernel's workgroup and tile (dgk_set_tile_shape): 0 built-in, 1 4 waves x 16
              elements, 2 8 waves x 32 elements, 3 row split (4 waves x 4 elements), 4 4 waves x 16
              elements with the operator in global memory, 5 row split over 8 waves x 8 elements,
              6, 7, 8 4 waves x 16 elements compiled for 3, 4 or 5 waves per SIMD, 9, 10, 11 as
              7, 8, 6 with the RK epilogue operands loaded late; shapes without a kernel for the case are
              skipped
  grid_pct    persistent grid as a percentage of the occupancy-derived grid: 50, 100, 200, 400
  tile_order  XCD-aware slabs (0) or round-robin (1)

The table is written as text lines "dim order precision variant grid_pct tile_order tile_shape
stage_us" (dgk_load_tuning) to dgfem-acoustic_amd/tuning.txt, which dgk.py loads with the
library. Results never depend on the launch shape and agree with the built-in tile shape to
rounding; the test suite checks both (tests/test_gpu_tuning.py).

Usage: python dgfem-acoustic_amd/autotune.py [--dofs 4e6] [--out PATH] [--dims 2,3] [--orders 1-5]
       python dgfem-acoustic_amd/autotune.py --time-only [--steps 5] [--dims 3] [--orders 4,5]
(--time-only: time the shipped table's choice on each shape, for profiling runs; writes nothing)
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
os.environ["DGK_NO_TUNING"] = "1"  # time the untuned defaults and the candidates explicitly

import dgk  # noqa: E402
from workloads import conf  # noqa: E402

GRID = (50, 100, 200, 400)
SHAPE_MARGIN = 0.98  # a non-built-in tile shape must beat the built-in one by 2%
ORDERS = (0, 1)


def mesh_for(dim, p, dofs):
    from math import comb

    Np = comb(p + dim, dim)
    K = max(64, int(dofs / ((dim + 1) * Np)))
    text = conf("Wall=Reflecting\nAbs=Absorbing\n")
    if dim == 1:
        return dgk.gen_line(K, 1.0, p, [1, 2], "1:Wall;2:Abs", text)
    if dim == 2:
        n = max(2, int(round(np.sqrt(K / 2))))
        return dgk.gen_rect(n, n, 1.0, 1.0, p, [1, 2, 1, 2], "1:Wall;2:Abs", text)
    n = max(2, int(round(np.cbrt(K / 6))))
    return dgk.gen_box(n, n, n, 1.0, 1.0, 1.0, p, 0.05, 1, [1, 2, 1, 2, 1, 2], "1:Wall;2:Abs", text)


def time_case(m, precision, variant, nsteps=10):
    """Stage time (us) of every candidate, keyed (variant, grid_pct, tile_order, tile_shape)."""
    s = dgk.Solver(m, 340.0, 1.225, precision=precision, flux_sign=dgk.SIGN_STABLE, variant=variant)
    u = np.zeros((4, m.K * m.Np))
    u[0] = np.exp(-((m.node_xyz - m.node_xyz.mean(axis=0)) ** 2).sum(axis=1) / 0.01)
    s.set_state(u)
    res = {}
    mfma = variant == dgk.VARIANT_MFMA
    for sh in dgk.TILE_SHAPES if mfma else (0,):
        if mfma:
            try:
                s.set_tile_shape(sh)
            except dgk.DGKError as e:
                if e.status != dgk.DGK_ERR_UNSUPPORTED:
                    raise
                continue  # no kernel of this shape for the case
        for g in GRID if mfma else (100,):
            for o in ORDERS if mfma else (0,):
                s.set_launch(g, o)
                s.rk4_timed(2, 1e-7)  # warm the launch shape
                ms, _ = s.rk4_timed(nsteps, 1e-7)
                res[(variant, g, o, sh)] = ms * 1e3 / (4 * nsteps)
    s.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--dofs", type=float, default=8e6)
    ap.add_argument("--dims", default="1,2,3")
    ap.add_argument("--orders", default="1,2,3,4,5")
    ap.add_argument("--precisions", default="fp64,fp32")
    ap.add_argument("--no-lds", action="store_true", help="skip the thread-per-node variant")
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.txt"))
    ap.add_argument("--time-only", action="store_true",
                    help="run `--steps` RK4 steps of each shape with the shipped tuning table's launch settings "
                         "(for counter runs) and print the stage time")
    ap.add_argument("--steps", type=int, default=5)
    args = ap.parse_args()
    if args.time_only:
        L = dgk.device()
        if L.dgk_load_tuning(dgk.TUNING_FILE.encode()):
            raise RuntimeError(L.dgk_last_error(None).decode())
        precs = {"fp64": dgk.FP64, "fp32": dgk.FP32}
        table = {}
        for line in open(dgk.TUNING_FILE):
            w = line.split()
            if w and not w[0].startswith("#"):
                table[(int(w[0]), int(w[1]), int(w[2]))] = w[3:]
        for dim in [int(d) for d in args.dims.split(",")]:
            for p in [int(o) for o in args.orders.split(",")]:
                m = mesh_for(dim, p, args.dofs)
                for pname in args.precisions.split(","):
                    s = dgk.Solver(m, 340.0, 1.225, precision=precs[pname], flux_sign=dgk.SIGN_STABLE)
                    u = np.zeros((4, m.K * m.Np))
                    u[0] = np.exp(-((m.node_xyz - m.node_xyz.mean(axis=0)) ** 2).sum(axis=1) / 0.01)
                    s.set_state(u)
                    ms, _ = s.rk4_timed(args.steps, 1e-7)
                    row = table.get((dim, p, precs[pname]))
                    print(f"[autotune] time-only dim={dim} p={p} {pname} K={m.K}: {ms * 1e3 / (4 * args.steps):.1f} us "
                          f"per stage (table row: {' '.join(row) if row else 'none'})", flush=True)
                    s.close()
        return
    lines = ["# dim order precision variant grid_pct tile_order tile_shape stage_us  (written by autotune.py on %s)"
             % time.strftime("%Y-%m-%d")]
    precs = {"fp64": dgk.FP64, "fp32": dgk.FP32}
    for dim in [int(d) for d in args.dims.split(",")]:
        for p in [int(o) for o in args.orders.split(",")]:
            m = mesh_for(dim, p, args.dofs)
            for pname in args.precisions.split(","):
                prec = precs[pname]
                res = time_case(m, prec, dgk.VARIANT_MFMA)
                # the thread-per-node kernel implements the FP64 form and the plain FP32 form only;
                # the default FP32 form (compensated solution) stays on the MFMA kernel
                if not args.no_lds and prec == dgk.FP64:
                    res.update(time_case(m, prec, dgk.VARIANT_LDS))
                (v, g, o, sh), us = min(res.items(), key=lambda kv: kv[1])
                # another tile shape replaces the built-in one only when it is clearly faster (timings
                # of one box repeat to about 1%; a different shape at the ensemble's sizes is not free)
                builtin = {k: t for k, t in res.items() if k[0] == dgk.VARIANT_MFMA and k[3] == 0}
                if sh != 0 and builtin:
                    k0, t0 = min(builtin.items(), key=lambda kv: kv[1])
                    if us > SHAPE_MARGIN * t0:
                        (v, g, o, sh), us = k0, t0
                base = res[(dgk.VARIANT_MFMA, 100, 0, 0)]
                shapes = {k[3]: min(t for kk, t in res.items() if kk[0] == dgk.VARIANT_MFMA and kk[3] == k[3])
                          for k in res if k[0] == dgk.VARIANT_MFMA}
                print(f"[autotune] dim={dim} p={p} {pname} K={m.K}: best variant={v} shape={sh} grid={g}% order={o} "
                      f"{us:.1f} us (default {base:.1f} us, {base / us:.2f}x; best per tile shape: "
                      + ", ".join(f"{k}: {t:.1f}" for k, t in sorted(shapes.items())) + ")", file=sys.stderr,
                      flush=True)
                lines.append(f"{dim} {p} {prec} {v} {g} {o} {sh} {us:.2f}")
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(args.out)


if __name__ == "__main__":
    main()
