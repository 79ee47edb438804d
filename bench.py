"""Benchmark: DOF-RK-stage updates/s of the DG hot path on one MI355X, FP64 plain vs FP32 batched.

Three legs, each timed over K RK4 steps after W warmup steps, with inputs resident in HBM:

* fp64_plain -- the drop-in path on the 3D order-3 case: unit cube, 32^3 cells Kuhn-split into
  196,608 tetrahedra, interior nodes jittered by 10% of h (seed 2), v0 = (50, 0, 0) m/s,
  absorbing walls, Gaussian pulse (sigma 0.04) at a seeded point, dt = 0.1 h_min/((c0+|v0|)(2p+1)).
* fp32_single -- the same case in the FP32 form (MFMA FP32, compensated FP32 + bf16 solution,
  FP64 geometry rounded once); relative L2 per field against the FP64 leg after the same steps.
* value: fp32 batched -- 256 independent simulations of uneven size (seeds 7+i): 128 2D
  rectangles and 128 3D Kuhn boxes with sides U[1,3], K ~ U[5k, 80k], order ~ U{2..5}, walls
  reflecting/absorbing with p = 1/2, Gaussian pulses with sigma ~ U[0.02, 0.08] x the shortest
  side; every simulation with its own CFL-0.1 step. One batch: groups of equal (dim, order) as
  parallel branches of one captured hipGraph of K steps. Its FP32 error is measured on 8
  members (every 32nd) against FP64 runs of the same members for the same steps.

A "step" is one RK4 step (4 stage launches) of every simulation in the leg. Counted unknowns
are (dim+1) fields per node. Synthetic data: meshes and states are generated here.

* c1_fp64 -- the reference's CPU-runnable case (2D unit square, 3200 triangles, order 3, 1000
  steps, FP64, reference flux sign), with the oracle's literal path timed on the host beside it.

* host_inclusive (optional leg `hostio`) -- the FP64 cube and the FP32 ensemble with the state
  handed in and out through host memory, as the C ABI's state calls take it, beside the same steps
  resident in HBM.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--legs fp64,fp32,batch,c1,hostio]
"""
import argparse
import json
import os
import sys
import time
from math import comb

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dgfem-acoustic_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import dgk  # noqa: E402
import workloads as wl  # noqa: E402

METRIC = "DOF-RK-stage updates/s: FP64 plain vs FP32-MFMA batched at stated L2 err"
UNIT = "DOF-RK-stage updates/s"
HBM_PEAK = 8.0e12
MFMA_PEAK = {4: 157.3e12, 8: 78.6e12}  # dense FP32 / FP64 matrix-core peaks of MI355X, flop/s
FIELDS = ["p", "vx", "vy", "vz"]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def element_sizes(dim, p):
    """(Np, Nf, Nfp) of the order-p simplex."""
    return comb(p + dim, dim), dim + 1, 1 if dim == 1 else comb(p + dim - 1, dim - 1)


def algorithmic_bytes(dim, p, K, F, Fb, word, compensated):
    """Algorithmic HBM bytes of one stage launch over K elements of one (dim, order).

    Per counted DOF ((dim+1) fields per node): the stage state traffic averaged over the four
    RK4 stages -- stage 1 reads u and writes the accumulator and the next stage input (3 words),
    stages 2-3 read the stage input, u and the accumulator and write the accumulator and the
    next input (5), stage 4 reads the stage input and the accumulator and writes u (3; 5 in the
    compensated FP32 form, which also reads u and reads/writes the 2-byte low word) -- plus the
    neighbour trace of every interior face read by each side. Per element: J^-1 (dim^2 words);
    per element face: unit normal and face scale (dim+1 words) and an int32 neighbour and int32
    permutation code.
    """
    Np, Nf, Nfp = element_sizes(dim, p)
    nf = dim + 1
    dof = nf * K * Np
    state = (16 + (2 if compensated else 0)) / 4 * word  # bytes per DOF per stage
    traces = 2 * (F - Fb) * Nfp * nf * word
    geo = K * (dim * dim * word + Nf * ((dim + 1) * word + 8))
    return dof * state + traces + geo


def algorithmic_flops(dim, p, K):
    """Matrix-core flops of one stage launch over K elements: per field and element the product
    [W_1..W_d | -LIFT] (Np x (d*Np + Nf*Nfp)) times that field's contravariant fluxes and face
    fluxes, i.e. 2*(d*Np + Nf*Nfp) flops per counted DOF ((dim+1) fields per node). The face flux,
    the flux contraction with J^-1 and the RK update run on the vector ALU and are not counted."""
    Np, Nf, Nfp = element_sizes(dim, p)
    return 2.0 * (dim * Np + Nf * Nfp) * (dim + 1) * K * Np


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the rocprofv3 FETCH_SIZE/WRITE_SIZE
    summaries committed under profiles/ (calibrated against a device copy there), or None."""
    d = os.path.join(ROOT, "profiles")
    key = kernel.split(">")[0].replace(" ", "")
    if not os.path.isdir(d):
        return None
    for f in sorted(os.listdir(d)):
        if not f.endswith(".md"):
            continue
        for line in open(os.path.join(d, f)):
            cols = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cols) == 8 and cols[0].replace(" ", "").startswith(key + ",") and cols[1] == workload:
                return {"bytes": int(float(cols[6])), "source": "profiles/" + f}
    return None


def measured_mfma_busy(kernel, workload):
    """Fraction of the matrix cores' cycles busy during `kernel` on `workload`, from the rocprofv3
    counter summaries under profiles/ (table rows "| kernel | workload | SQ_VALU_MFMA_BUSY_CYCLES |
    GRBM_GUI_ACTIVE | CUs | busy |", busy = MFMA busy cycles / (GUI-active cycles x 4 SIMDs x CUs)),
    or None."""
    d = os.path.join(ROOT, "profiles")
    key = kernel.split(">")[0].replace(" ", "")
    if not os.path.isdir(d):
        return None
    for f in sorted(os.listdir(d)):
        if not f.endswith(".md"):
            continue
        for line in open(os.path.join(d, f)):
            cols = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cols) == 6 and cols[0].replace(" ", "").startswith(key + ",") and cols[1] == workload:
                try:
                    return {"frac": float(cols[5]), "source": "profiles/" + f}
                except ValueError:
                    continue
    return None


def roofline(bytes_per_launch, flops_per_launch, word, stage_s, kernel, workload):
    """The bound is the roof the kernel's algorithmic intensity (flops per HBM byte) meets first:
    HBM below the ridge MFMA_PEAK/HBM_PEAK, the matrix cores above it."""
    peak_f = MFMA_PEAK[word]
    mfma = flops_per_launch / bytes_per_launch > peak_f / HBM_PEAK
    ach_b, ach_f = bytes_per_launch / stage_s, flops_per_launch / stage_s
    tr = measured_traffic(kernel, workload)
    if mfma:
        out = {"bound": "mfma", "achieved": round(ach_f / 1e12, 2), "peak": peak_f / 1e12, "unit": "TFLOP/s",
               "frac": round(ach_f / peak_f, 4)}
    else:
        out = {"bound": "hbm", "achieved": round(ach_b / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
               "frac": round(ach_b / HBM_PEAK, 4)}
    out.update({"traffic": tr["bytes"] if tr else None, "kernel": kernel, "kernel_avg_us": round(stage_s * 1e6, 2),
                "algorithmic_bytes_per_launch": int(bytes_per_launch),
                "algorithmic_flops_per_launch": int(flops_per_launch),
                "intensity_flop_per_byte": round(flops_per_launch / bytes_per_launch, 2),
                "ridge_flop_per_byte": round(peak_f / HBM_PEAK, 2),
                "algorithmic_GBps": round(ach_b / 1e9, 1), "mfma_TFLOPs": round(ach_f / 1e12, 2)})
    if tr:
        out["traffic_source"] = tr["source"]
        out["traffic_GBps"] = round(tr["bytes"] / stage_s / 1e9, 1)
    mb = measured_mfma_busy(kernel, workload)
    out["mfma_busy"] = mb["frac"] if mb else None
    if mb:
        out["mfma_busy_source"] = mb["source"]
    return out


def single_leg(m, ph, dt, u0, precision, steps, warmup):
    s = dgk.Solver(m, ph["c0"], ph["rho0"], v0=ph["v0"], precision=precision, flux_sign=dgk.SIGN_STABLE)
    s.set_state(u0)
    s.rk4(warmup, dt)
    s.sync()
    t0 = time.perf_counter()
    s.rk4(steps, dt)
    s.sync()
    wall = time.perf_counter() - t0
    nk = 50  # per-launch kernel time: HIP events around every stage launch on the context's stream
    stage_ms, _ = s.rk4_timed(nk, dt)
    state = s.get_state()
    info = s.info
    s.close()
    dof = (m.dim + 1) * m.K * m.Np
    word = 8 if precision == dgk.FP64 else 4
    b = algorithmic_bytes(m.dim, m.order, m.K, info.num_faces, info.num_boundary_faces, word, precision == dgk.FP32)
    stage_s = stage_ms * 1e-3 / (4 * nk)
    return dict(rate=dof * 4 * steps / wall, wall=wall, state=state, total_steps=warmup + steps + nk,
                roofline=roofline(b, algorithmic_flops(m.dim, m.order, m.K), word, stage_s,
                                  "k_stage_mfma<%s,3,3>" % ("double" if word == 8 else "float"), "cube3d-p3"),
                finite=bool(np.isfinite(state).all()))


def batch_leg(steps, warmup, members_checked=8, n=256):
    t0 = time.perf_counter()
    meshes, phys, dts, pulses = wl.ensemble(n=n)
    log(f"batch: generated {n} meshes in {time.perf_counter() - t0:.1f} s, "
        f"K_total={sum(m.K for m in meshes)}, DOF={wl.dof_count(meshes)}")
    t0 = time.perf_counter()
    b = dgk.Batch(meshes, phys, dts, precision=dgk.FP32, flux_sign=dgk.SIGN_STABLE)
    for i, (m, pu) in enumerate(zip(meshes, pulses)):
        b.set_state(i, wl.ensemble_state(m, pu))
    groups = b.groups()
    log(f"batch: device setup and state upload {time.perf_counter() - t0:.1f} s, {len(groups)} groups")
    b.rk4(warmup)
    # tile shape and launch shape of every group at its own size, then the schedule: groups as
    # parallel graph branches or in sequence, whichever runs a short captured trial faster on this
    # GPU (10 steps each)
    t0 = time.perf_counter()
    tuned, tune_steps = b.tune_groups()
    log(f"batch: group tuning ({tune_steps} event-timed steps, {time.perf_counter() - t0:.1f} s): " +
        ", ".join(f"{g['dim']}D p{g['order']} shape {c[0]} grid {c[1]} order {c[2]} ({c[3]:.1f} us)"
                  for g, c in zip(groups, tuned)))
    schedule, trial = b.pick_schedule(10)
    log(f"batch: schedule trial (10 steps) parallel {trial[0]:.3f} s, sequential {trial[1]:.3f} s")
    total_before = warmup + tune_steps + 20
    t0 = time.perf_counter()
    b.prepare(steps, schedule)
    log(f"batch: captured {steps} steps x {len(groups)} groups in {time.perf_counter() - t0:.1f} s")
    b.sync()
    t0 = time.perf_counter()
    b.rk4(steps, schedule)
    b.sync()
    wall = time.perf_counter() - t0
    dof = wl.dof_count(meshes)
    nk = 10
    gms = b.rk4_timed(nk)
    total = total_before + steps + nk
    gtab = []
    for g, ms, c in zip(groups, gms, tuned):
        Np = element_sizes(g["dim"], g["order"])[0]
        bl = algorithmic_bytes(g["dim"], g["order"], g["elements"], g["faces"], g["boundary_faces"], 4, True)
        st = ms * 1e-3 / (4 * nk)
        gtab.append(dict(dim=g["dim"], order=g["order"], sims=g["sims"], elements=g["elements"],
                         tile_shape=c[0], grid_pct=c[1], tile_order=c[2],
                         dof=(g["dim"] + 1) * g["elements"] * Np, stage_us=round(st * 1e6, 1), bytes=bl,
                         flops=algorithmic_flops(g["dim"], g["order"], g["elements"]), stage_s=st))
    tot = sum(x["stage_s"] for x in gtab)
    for x in gtab:
        x["share"] = round(x["stage_s"] / tot, 3)
        x["GBps"] = round(x["bytes"] / x["stage_s"] / 1e9, 1)
        x["TFLOPs"] = round(x["flops"] / x["stage_s"] / 1e12, 2)
    dom = max(gtab, key=lambda x: x["stage_s"])
    roof = roofline(dom["bytes"], dom["flops"], 4, dom["stage_s"],
                    "k_stage_mfma<float,%d,%d> (largest batch group)" % (dom["dim"], dom["order"]), "ensemble256")
    roof["sum_of_group_stage_us"] = round(tot * 1e6, 1)
    roof["all_groups_GBps"] = round(sum(x["bytes"] for x in gtab) / tot / 1e9, 1)
    # FP32 error on a subset against FP64 runs of the same members
    chk = list(range(0, n, n // members_checked))
    g32 = {i: b.get_state(i) for i in chk}
    b.close()
    sub = dgk.Batch([meshes[i] for i in chk], [phys[i] for i in chk], [dts[i] for i in chk], precision=dgk.FP64,
                    flux_sign=dgk.SIGN_STABLE)
    for j, i in enumerate(chk):
        sub.set_state(j, wl.ensemble_state(meshes[i], pulses[i]))
    sub.rk4(total)
    errs = {f: 0.0 for f in FIELDS}
    finite = True
    for j, i in enumerate(chk):
        r = sub.get_state(j)
        finite &= bool(np.isfinite(g32[i]).all())
        for q in range(meshes[i].dim + 1):
            errs[FIELDS[q]] = max(errs[FIELDS[q]], float(np.linalg.norm(g32[i][q] - r[q]) / np.linalg.norm(r[q])))
    sub.close()
    for x in gtab:
        del x["stage_s"], x["bytes"], x["flops"]
    return dict(rate=dof * 4 * steps / wall, wall=wall, dof=dof, elements=sum(m.K for m in meshes), groups=gtab,
                roofline=roof, errs=errs, checked=chk, total_steps=total, finite=finite,
                schedule="parallel" if schedule == dgk.SCHEDULE_PARALLEL else "sequential",
                sample=(meshes, phys, dts, pulses))


def small_leg(no_cpu, steps=1000):
    """The reference's CPU-runnable case: 2D unit square, 40x40 cells (3200 triangles), order 3,
    reflecting walls, Gaussian pulse, 1000 RK4 steps of FP64 on the GPU (drop-in path, reference
    flux sign), beside the oracle's literal path on the host over the first 50 steps."""
    m, ph, dt, u0 = wl.square_2d()
    dof = 3 * m.K * m.Np
    s = dgk.Solver(m, ph["c0"], ph["rho0"])
    s.set_state(u0)
    s.rk4(25, dt)
    s.set_state(u0)
    s.sync()
    t0 = time.perf_counter()
    s.rk4(steps, dt)
    s.sync()
    wall = time.perf_counter() - t0
    g = s.get_state()
    s.close()
    out = {"workload": "square2d-p3: unit square, 40x40 cells x 2 triangles (3200), order 3, reflecting walls, "
                       "Gaussian pulse, FP64, reference flux sign", "steps": steps, "value": dof * 4 * steps / wall,
           "ms_per_step": wall * 1e3 / steps, "unit": UNIT}
    log(f"c1: {out['value']:.4g} DOF-stage/s ({wall * 1e3:.1f} ms for {steps} steps)")
    if not no_cpu:
        import oracle_binding as ob

        threads = os.cpu_count() or 1
        cpu_steps = 50  # bounded sample of the same run
        o = ob.Oracle(m, ph["c0"], ph["rho0"], nthreads=threads)
        t0 = time.perf_counter()
        ref = o.step(u0, cpu_steps, dt, mode=ob.LITERAL)
        sec = time.perf_counter() - t0
        s = dgk.Solver(m, ph["c0"], ph["rho0"])
        s.set_state(u0)
        s.rk4(cpu_steps, dt)
        g = s.get_state()
        s.close()
        out["cpu_literal"] = {"value": dof * 4 * cpu_steps / sec, "sample": f"first {cpu_steps} steps, {sec:.1f} s",
                              "cores": threads, "kind": "port"}
        out["rel_l2_vs_cpu_after_50_steps"] = float(np.linalg.norm(g - ref) / np.linalg.norm(ref))
        log(f"c1: oracle literal {sec:.1f} s for {cpu_steps} steps on {threads} threads, GPU vs CPU rel L2 "
            f"{out['rel_l2_vs_cpu_after_50_steps']:.2e}")
    return out


def host_cpu():
    """Model name and logical CPU count of the host (from /proc/cpuinfo)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def smallest_member_per_shape(n=256):
    """Index of the smallest member (by drawn element count) of every (dim, order) of the ensemble."""
    best = {}
    for i in range(n):
        dim, K, p = wl.ensemble_draw(i, n=n)[:3]
        if (dim, p) not in best or K < best[(dim, p)][0]:
            best[(dim, p)] = (K, i)
    return sorted(i for _, i in best.values())


def cpu_baseline(meshes, steps=6):
    """The oracle on a bounded sample of the ensemble, all host threads: `steps` RK4 steps of one
    member per (dim, order) -- the smallest of its shape, drawn again with about 3000 (2D), 1500
    (3D orders 2-3) or 600 (3D orders 4-5) elements (same seed, shape, sides, walls and pulse), about
    10 s of host time in all -- with its literal path (the reference's loop
    structure, quadrature-point arrays and OpenMP layout) and its factorized path. The value is the
    rate at which the whole ensemble would run: the per-(dim, order) rates weighted by the
    ensemble's unknowns of that shape (total unknown-stages over the summed time they would take)."""
    import oracle_binding as ob

    threads = os.cpu_count() or 1
    members = smallest_member_per_shape()
    sm, sp, sd, su = [], [], [], []
    for i in members:
        dim, _, p = wl.ensemble_draw(i)[:3]
        k = 3000 if dim == 2 else 1500 if p <= 3 else 600
        for a, b in zip((sm, sp, sd, su), wl.ensemble(n=256, members=[i], k_target=k, threads=1)):
            a.extend(b)
    rates = {}
    t_lit = 0.0
    for m, ph, dt, pu in zip(sm, sp, sd, su):
        u = wl.ensemble_state(m, pu)
        o = ob.Oracle(m, ph["c0"], ph["rho0"], v0=ph["v0"], sign_mode=ob.SIGN_STABLE, nthreads=threads)
        t0 = time.perf_counter()
        o.step(u, steps, dt, mode=ob.LITERAL)
        tl = time.perf_counter() - t0
        t0 = time.perf_counter()
        o.step(u, steps, dt, mode=ob.FACTORIZED)
        tf = time.perf_counter() - t0
        del o
        w = (m.dim + 1) * m.K * m.Np * steps * 4
        rates[(m.dim, m.order)] = (w / tl, w / tf)
        t_lit += tl
    share = {}
    for m in meshes:
        share[(m.dim, m.order)] = share.get((m.dim, m.order), 0) + (m.dim + 1) * m.K * m.Np
    tot = sum(share.values())
    lit = tot / sum(share[k] / rates[k][0] for k in share)
    fac = tot / sum(share[k] / rates[k][1] for k in share)
    desc = ", ".join(f"member {i} ({m.dim}D p{m.order} K={m.K})" for i, m in zip(members, sm))
    return {"value": lit, "unit": UNIT, "cores": threads, "kind": "port", "host_cpu": host_cpu(),
            "sample": f"{steps} RK4 steps each of {desc} (one member per dim/order) with the oracle's literal "
                      f"path (the reference's loop structure and OpenMP layout), {t_lit:.1f} s; value: the "
                      f"per-shape rates weighted by the ensemble's unknowns per shape; factorized_value: the "
                      f"oracle's factorized path, same weighting",
            "factorized_value": fac,
            "per_shape": {f"{d}D p{p}": {"literal": round(r[0], 1), "factorized": round(r[1], 1)}
                          for (d, p), r in sorted(rates.items())}}


def cube_cpu(steps=8, n=12):
    """The oracle beside the FP64 cube leg: `steps` RK4 steps of the same generator at n^3 cells
    (the full 32^3 cube takes about a minute per step on the literal path), literal and factorized
    paths, all host threads."""
    import oracle_binding as ob

    threads = os.cpu_count() or 1
    m, ph, dt, u0 = wl.cube_3d(n=n)
    o = ob.Oracle(m, ph["c0"], ph["rho0"], v0=ph["v0"], sign_mode=ob.SIGN_STABLE, nthreads=threads)
    w = (m.dim + 1) * m.K * m.Np * steps * 4
    t0 = time.perf_counter()
    o.step(u0, steps, dt, mode=ob.LITERAL)
    tl = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.step(u0, steps, dt, mode=ob.FACTORIZED)
    tf = time.perf_counter() - t0
    del o
    return {"value": w / tl, "unit": UNIT, "cores": threads, "kind": "port", "host_cpu": host_cpu(),
            "sample": f"first {steps} RK4 steps of the cube generator at {n}^3 cells ({m.K} tets, order 3), "
                      f"literal path {tl:.1f} s",
            "factorized_value": w / tf}


def host_io_leg(steps):
    """Rates with the state handed over in host memory, as the C ABI takes and returns it
    (`dgk_set_state` / `dgk_get_state` and their batch forms: FP64 host arrays, 4 field planes):
    host state in, `steps` RK4 steps, host state out, each phase timed on its own after a warm
    run, beside the same steps with the state resident in HBM (`value`'s form). The FP64 cube (the
    drop-in form) and the FP32 ensemble."""
    out = {}
    m, ph, dt, u0 = wl.cube_3d()
    dof = 4 * m.K * m.Np
    s = dgk.Solver(m, ph["c0"], ph["rho0"], v0=ph["v0"], precision=dgk.FP64, flux_sign=dgk.SIGN_STABLE)
    s.set_state(u0)
    s.rk4(min(steps, 25), dt)
    s.get_state()
    t = {}
    s.sync()
    t0 = time.perf_counter()
    s.set_state(u0)
    s.sync()
    t["in"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    s.rk4(steps, dt)
    s.sync()
    t["run"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    s.get_state()
    t["out"] = time.perf_counter() - t0
    s.close()
    out["cube_fp64"] = {"steps": steps, "host_bytes_each_way": int(u0.nbytes), "host_in_s": t["in"],
                        "resident_s": t["run"], "host_out_s": t["out"], "resident_rate": dof * 4 * steps / t["run"],
                        "host_inclusive_rate": dof * 4 * steps / (t["in"] + t["run"] + t["out"])}
    log(f"hostio cube fp64: {out['cube_fp64']['host_inclusive_rate']:.4g} with host state in and out, "
        f"{out['cube_fp64']['resident_rate']:.4g} resident")
    meshes, phys, dts, pulses = wl.ensemble(n=256)
    states = [wl.ensemble_state(mm, pu) for mm, pu in zip(meshes, pulses)]
    dof = wl.dof_count(meshes)
    b = dgk.Batch(meshes, phys, dts, precision=dgk.FP32, flux_sign=dgk.SIGN_STABLE)
    for i, u in enumerate(states):
        b.set_state(i, u)
    schedule, _ = b.pick_schedule(10)
    b.rk4(min(steps, 100), schedule)
    for i in range(len(states)):
        b.get_state(i)
    b.sync()
    t0 = time.perf_counter()
    for i, u in enumerate(states):
        b.set_state(i, u)
    b.sync()
    t["in"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.rk4(steps, schedule)
    b.sync()
    t["run"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    for i in range(len(states)):
        b.get_state(i)
    t["out"] = time.perf_counter() - t0
    b.close()
    out["ensemble_fp32"] = {"steps": steps, "host_bytes_each_way": int(sum(u.nbytes for u in states)),
                            "host_in_s": t["in"], "resident_s": t["run"], "host_out_s": t["out"],
                            "resident_rate": dof * 4 * steps / t["run"],
                            "host_inclusive_rate": dof * 4 * steps / (t["in"] + t["run"] + t["out"])}
    log(f"hostio ensemble fp32: {out['ensemble_fp32']['host_inclusive_rate']:.4g} with host state in and out, "
        f"{out['ensemble_fp32']['resident_rate']:.4g} resident")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--legs", default="fp64,fp32,batch,c1",
                    help="comma list of fp64,fp32,batch,c1,hostio (hostio: rates with the state handed over "
                         "in host memory; not in the default run)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="first run a plain device copy of MB MiB (reference dispatch for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the sharding design is documented, "
              "not built)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calibration} MiB read + {args.copy_calibration} MiB written")
        del src, dst
    legs = args.legs.split(",")
    res = {}
    if "fp64" in legs or "fp32" in legs:
        m, ph, dt, u0 = wl.cube_3d()
        for name, prec in (("fp64", dgk.FP64), ("fp32", dgk.FP32)):
            if name not in legs:
                continue
            log(f"{name} cube: {args.warmup} warmup + {args.steps} timed RK4 steps, K={m.K}")
            res[name] = single_leg(m, ph, dt, u0, prec, args.steps, args.warmup)
            log(f"{name} cube: {res[name]['rate']:.4g} DOF-stage/s, stage kernel "
                f"{res[name]['roofline']['kernel_avg_us']} us ({res[name]['roofline']['achieved']} GB/s)")
        if "fp64" in legs and not args.no_cpu:
            res["fp64"]["cpu"] = cube_cpu()
            log(f"fp64 cube: oracle literal {res['fp64']['cpu']['value']:.4g}, factorized "
                f"{res['fp64']['cpu']['factorized_value']:.4g} DOF-stage/s on {res['fp64']['cpu']['cores']} threads")
    if "c1" in legs:
        res["c1"] = small_leg(args.no_cpu)
    if "batch" in legs:
        log(f"batch: {args.warmup} warmup + {args.steps} timed RK4 steps")
        res["batch"] = batch_leg(args.steps, args.warmup)
        log(f"batch: {res['batch']['rate']:.4g} DOF-stage/s, {res['batch']['wall']:.1f} s")
    out = {"metric": METRIC, "unit": UNIT, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "data": "synthetic"}
    if "fp32" in res and "fp64" in res:
        g64, g32 = res["fp64"]["state"], res["fp32"]["state"]
        res["fp32"]["errs"] = {FIELDS[q]: float(np.linalg.norm(g32[q] - g64[q]) / np.linalg.norm(g64[q]))
                               for q in range(4)}
    cube_cfg = ("cube3d-p3: unit cube, 32^3 Kuhn cells (196,608 tets), 10% jitter, order 3, 4 fields, "
                "v0=(50,0,0), absorbing walls, Gaussian pulse, flux sign stable")
    if "batch" in res:
        bt = res["batch"]
        out.update(value=bt["rate"], ms_per_step=bt["wall"] * 1e3 / args.steps, dtype="fp32",
                   config={"workload": "ensemble256: 256 uneven simulations (128 2D tri rectangles, 128 3D Kuhn "
                                       "boxes, K~U[5k,80k], order~U{2..5}, mixed walls, own dt each), FP32 "
                                       "compensated form, groups as branches of one captured hipGraph",
                           "simulations": 256, "elements": bt["elements"], "dofs": bt["dof"],
                           "flux_sign": "stable", "schedule": bt["schedule"]},
                   fp32_rel_l2_vs_fp64={"after_steps": bt["total_steps"], "members": bt["checked"], **bt["errs"]},
                   roofline=bt["roofline"], groups=bt["groups"], finite=bt["finite"])
    elif "fp32" in res:
        r = res["fp32"]
        out.update(value=r["rate"], ms_per_step=r["wall"] * 1e3 / args.steps, dtype="fp32",
                   config={"workload": cube_cfg}, roofline=r["roofline"])
    elif "fp64" in res:
        r = res["fp64"]
        out.update(value=r["rate"], ms_per_step=r["wall"] * 1e3 / args.steps, dtype="fp64",
                   config={"workload": cube_cfg}, roofline=r["roofline"])
    if "fp64" in res:
        r = res["fp64"]
        out["fp64_plain"] = {"value": r["rate"], "ms_per_step": r["wall"] * 1e3 / args.steps, "dtype": "fp64",
                             "workload": cube_cfg, "roofline": r["roofline"], "finite": r["finite"],
                             "cpu_baseline": r.get("cpu")}
        if "batch" in res:
            out["fp32_batched_over_fp64_plain"] = res["batch"]["rate"] / r["rate"]
    if "fp32" in res:
        r = res["fp32"]
        out["fp32_single"] = {"value": r["rate"], "ms_per_step": r["wall"] * 1e3 / args.steps, "dtype": "fp32",
                              "workload": cube_cfg, "roofline": r["roofline"], "finite": r["finite"]}
        if "errs" in r:
            out["fp32_single"]["rel_l2_vs_fp64"] = {"after_steps": r["total_steps"], **r["errs"]}
    if "c1" in res:
        out["c1_fp64"] = res["c1"]
    if "hostio" in legs:
        out["host_inclusive"] = host_io_leg(args.steps)
    out["cpu_baseline"] = None
    if not args.no_cpu and "batch" in res:
        log("cpu baseline: literal oracle on a sample of the ensemble, all host threads")
        out["cpu_baseline"] = cpu_baseline(res["batch"]["sample"][0])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
