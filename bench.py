#!/usr/bin/env python3
"""Throughput of the Biot time loop on one MI355X.

Workload (the roofline case): layered half-space 8000 x 4000 cells, dx = 2 m,
10 seeded layers times a seeded Gaussian-correlated field (50 m, 5 % std) on
porosity, Kdry and mu (quantized to 16 levels per parameter, so the material
table stays indexed), free surface by stress imaging, 20-cell PML on the other
sides, 20 Hz Ricker source at the surface, receivers on the surface. A "step"
is one DO ITILE iteration of TIME_LOOP: stress half-step, velocity half-step,
edge rows and receiver samples, with all inputs resident in HBM.

Prints one JSON line. ``--gpus N`` with N > 1 is refused: this build runs on
one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "2d-vqsp-fd-sim_amd"))

from biotfd import formats, hostsetup, lib, models  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of one cell (indexed materials): each half-step reads the
# other group's 4 fields, reads and writes its own 4 and reads the int32
# material index, 52 B; a whole step (both half-steps) is 104 B.
B_STEP = 104
B_HALF = 52
# What the fused interior kernel must move at least for one cell-step: 8 fields
# read once, 8 written once, 1 index (the other group's fields are consumed in
# registers and LDS within the same sweep). The index is stored in 16 bits
# where the table has at most 65535 materials.
B_FUSED_MIN = 68
B_FUSED_MIN_IDX16 = 66
# Per-cell materials (every cell its own material): the index is replaced by
# the 22 precombined coefficients a cell's update needs, 10 for the stress
# half-step and 12 for the velocity half-step: 88 + 96 = 184 B per cell-step.
B_STEP_CELL = 184
B_HALF_CELL = 88
B_FUSED_MIN_CELL = 64 + 88
TIMING_STRIDE = 8     # kernel durations from HIP events around every 8th step
TUNE_TRIAL = 4        # --tiles tune: timed steps per bq_autotune candidate


def make_workload(steps: int, mx=8000, mz=4000, per_cell=False):
    """The roofline model: 10 seeded layers times a Gaussian-correlated 5 %
    field on porosity, Kdry and mu, quantized to 16 levels per parameter
    (indexed materials) or, with per_cell, one material per cell."""
    job = models.layered_job(name="ROOF", mx=mx, mz=mz, h=2.0, steps=steps, n_layers=10, seed=2,
                             tpml=20, f0=20.0, n_rec=1000, field_std=0.05, corr_len=50.0,
                             levels=None if per_cell else 16)
    return job


def cpu_reference_baseline(job_full, steps_a=2, steps_b=12, budget_s=300.0):
    """Times the reference (built from its sources into oracle/_ref/fd2d, serial,
    one core) on the same grid: per-step time from two runs that differ only in
    the step count, so set-up and input reading cancel."""
    exe = os.path.join(ROOT, "oracle", "_ref", "fd2d")
    if not os.path.exists(exe):
        return None
    times = {}
    tmp = tempfile.mkdtemp(prefix="bqref_")
    try:
        for n in (steps_a, steps_b):
            job = make_like(job_full, n)
            d = os.path.join(tmp, f"s{n}")
            formats.write_job(job, d)
            t0 = time.perf_counter()
            subprocess.run([exe], cwd=d, check=True, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL, timeout=budget_s)
            times[n] = time.perf_counter() - t0
            print(f"[bench] reference {n} steps: {times[n]:.2f} s", file=sys.stderr, flush=True)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    per_step = (times[steps_b] - times[steps_a]) / (steps_b - steps_a)
    return per_step


def cpu_port_baseline(job, steps=3):
    """The CPU baseline on all host threads: the binary32 oracle (a C
    restatement of the reference's serial loop, OpenMP over rows, built with
    -O2 and no contraction) on the same grid; per step over `steps` steps
    after one untimed step. Returns (Gcell/s, threads)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ref = oracle.OracleSim(job, "f32")
    ref.run(1)
    t0 = time.perf_counter()
    ref.run(steps)
    dt = (time.perf_counter() - t0) / steps
    ref.close()
    nz, nx = hostsetup.grid_shape(job)
    print(f"[bench] oracle on {threads} threads: {dt:.2f} s per step", file=sys.stderr, flush=True)
    return nx * nz / dt / 1e9, threads


def make_like(job, steps):
    import copy
    j = copy.copy(job)
    j.mt = float(np.float32((steps - 0.5) * np.float32(job.dt)))
    j.key_tlv = False
    j.key_snv = True
    j.key_tlp = False
    j.ipas2 = 10 ** 8
    j.receivers = []
    assert j.mt2 == steps
    return j


def _digest(*arrays) -> str:
    import hashlib
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def run_batched(args):
    """Batched shots: N seeded layered half-spaces of sizes drawn from 500..3000 x
    500..1500 cells, 20-cell PML, surface source with f0 ~ U(10, 40) Hz, up to
    1000 surface receivers, warmup + steps time steps each. First every shot
    alone in sequence (bq_run, the plain form), then all of them as one batch
    (bq_batch_run), whose warmup steps are the autotuner's trials. Both time
    exactly `steps` steps per shot with inputs resident; the batch's fields and
    traces must equal the plain runs' bit for bit."""
    total = args.warmup + args.steps
    trial = max(1, args.warmup // 6)
    if args.warmup < 6:
        raise SystemExit("--warmup must be >= 6 for the batched workload (autotuner trials)")
    t0 = time.perf_counter()
    jobs = models.shot_set(args.shots, total, seed=1000)
    cells = [int(np.prod(hostsetup.grid_shape(j))) for j in jobs]
    print(f"[bench] {len(jobs)} shots ready, {sum(cells) / 1e6:.1f} M cells per step "
          f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    # plain: one shot after another, each with its own three launches per step
    per_shot, digests = [], []
    for j in jobs:
        sim = lib.Simulation(j)
        sim.run(args.warmup)
        a = time.perf_counter()
        sim.run(args.steps)   # returns after bq_sync
        per_shot.append(time.perf_counter() - a)
        digests.append(_digest(sim.traces(), sim.field("U"), sim.field("P")))
        sim.close()
    t_plain = sum(per_shot)
    print(f"[bench] plain: {t_plain:.2f} s for {len(jobs)} shots", file=sys.stderr, flush=True)
    # batched
    sims = [lib.Simulation(j) for j in jobs]
    b = lib.Batch(sims)
    tune = b.autotune(trial)
    if 6 * trial < args.warmup:
        b.run(args.warmup - 6 * trial)
    b.sync()
    a = time.perf_counter()
    b.run(args.steps, sync=False)
    b.sync()
    t_batch = time.perf_counter() - a
    dev_ms = b.last_run_ms()
    rows, sched, graph = b.schedule()
    same = all(_digest(s.traces(), s.field("U"), s.field("P")) == d for s, d in zip(sims, digests))
    b.close()
    for s in sims:
        s.close()
    if not same:
        raise SystemExit("batched results differ from the shots run alone")
    n = len(jobs)
    step_cells = sum(cells)
    gcell = step_cells * args.steps / t_batch / 1e9
    out = {
        "metric": "shots/s batched vs plain at <=1e-3 L2 misfit",
        "value": round(n / t_batch, 3),
        "unit": "shots/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_batch * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{n} batched shots, layered half-spaces 500..3000 x 500..1500 "
                               f"cells, dx=2 m, TPML=20, f0 U(10,40) Hz, surface receivers",
                   "shots": n, "cells_per_step": step_cells, "receivers": sum(len(j.receivers) for j in jobs)},
        "plain_shots_per_s": round(n / t_plain, 3),
        "speedup_vs_plain": round(t_plain / t_batch, 3),
        "plain_shot_s": {"p50": round(float(np.percentile(per_shot, 50)), 4),
                         "p99": round(float(np.percentile(per_shot, 99)), 4),
                         "max": round(max(per_shot), 4)},
        "batched_shot_s": {"p50": round(t_batch, 4), "p99": round(t_batch, 4),
                           "note": "every shot of the batch completes with the batch"},
        "l2_misfit_vs_plain": 0.0, "bit_identical_to_plain": True,
        "batch_device_ms": round(dev_ms, 2),
        "schedule": {"tile_rows": rows, "schedule": ["grid", "persistent"][sched], "hip_graph": graph},
        "autotune_ms_per_step": {f"{r}/{['grid', 'persistent'][sc]}": round(v, 4)
                                 for (r, sc), v in tune.items()},
        "gcell_per_s": round(gcell, 2),
        "roofline": {"bound": "hbm", "kernel": "whole batched step (kb_step, kb_rest, kb_velocity with the edge rows)",
                     "basis": "104 B per cell-step over all shots' cells",
                     "algorithmic_bytes": int(B_STEP * step_cells),
                     "achieved": round(gcell * B_STEP, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gcell * B_STEP / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        # the reference, serial, on the median-sized shot: per-step time from a
        # 12- and a 2-step run, times the steps of a shot
        med = jobs[int(np.argsort(cells)[len(cells) // 2])]
        per_step = cpu_reference_baseline(med)
        if per_step and per_step > 0:
            out["cpu_baseline"] = {
                "value": round(1.0 / (per_step * args.steps), 6), "unit": "shots/s", "cores": 1,
                "kind": "reference",
                "sample": f"reference fd2d (serial) on the median shot {med.mx}x{med.mz}, 12 minus "
                          f"2 steps, scaled to {args.steps} steps per shot"}
    print(json.dumps(out), flush=True)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_learned(args):
    """Learned absorbing boundary plus batching, on held-out shots drawn as in
    the batched workload (seeds from 20000). Every shot runs four ways: the
    40-cell PML, batched (the judge); the 8-cell strip with the trained
    network, batched; the plain 8-cell PML, batched; the 20-cell PML alone,
    one shot after another (the plain GPU form). A shot's learned run is
    accepted where its reflected boundary energy (U and W over the physical
    cells at every 500th step, against the judge) stays <= 1e-3 and its
    receiver traces' relative L2 misfit against the judge stays <= 1e-3;
    the others fall back to the plain 20-cell PML, rerun in one batch whose
    time is charged to the learned path."""
    from biotfd import learned
    total = args.steps
    net = learned.Network.load()
    if args.lb_gain is not None or args.lb_damping is not None:
        net = net.with_gain(net.gain if args.lb_gain is None else args.lb_gain,
                            None if args.lb_damping is None else bool(args.lb_damping))
    every = 500
    n = args.shots
    jobs = {t: models.shot_set(n, total, seed=20_000, tpml=t) for t in (40, 8, 20)}
    cells = {t: [int(np.prod(hostsetup.grid_shape(j))) for j in js] for t, js in jobs.items()}
    log(f"[bench] {n} held-out shots; cells per step: " +
        ", ".join(f"TPML {t}: {sum(c) / 1e6:.1f} M" for t, c in cells.items()))
    sims = {k: [lib.Simulation(j) for j in jobs[t]] for k, t in (("wide", 40), ("plain8", 8))}
    # the network's strip: the 8-cell CPML or the inner nodes of the judge's profile
    sims["learned"] = [net.strip_grid(j, w) for j, w in zip(jobs[8], sims["wide"])]
    for s in sims["learned"]:
        net.attach(s)
    batches = {k: lib.Batch(v) for k, v in sims.items()}
    for b in batches.values():
        b.set_schedule(256, lib.SCHED_GRID, True)
    dev_ms = {k: 0.0 for k in batches}
    wall = {k: 0.0 for k in batches}
    refl = {k: np.zeros((n, 2)) for k in ("learned", "plain8")}
    done = 0
    while done < total:
        m = min(every, total - done)
        for k, b in batches.items():
            a = time.perf_counter()
            b.run(m)          # returns after bq_batch_sync
            wall[k] += time.perf_counter() - a
            dev_ms[k] += b.last_run_ms()
        done += m
        for k in refl:
            for i in range(n):
                refl[k][i] += learned.reflected_energy(sims[k][i], sims["wide"][i])
        log(f"[bench] step {done}")
    with np.errstate(invalid="ignore", divide="ignore"):
        ratio = {k: v[:, 0] / v[:, 1] for k, v in refl.items()}

    def misfit(a, w):
        t = lib.trace_misfit(a, w)
        return float(np.sqrt(t[0] / t[1]))

    mis = {k: np.array([misfit(a, w) for a, w in zip(sims[k], sims["wide"])]) for k in refl}
    # the plain GPU form: 20-cell PML, each shot alone
    per_shot, m20 = [], []
    for j, w in zip(jobs[20], sims["wide"]):
        s = lib.Simulation(j)
        a = time.perf_counter()
        s.run(total)
        per_shot.append(time.perf_counter() - a)
        m20.append(misfit(s, w))
        s.close()
    mis["plain20"] = np.array(m20)
    log("[bench] plain GPU form done")
    t_plain = sum(per_shot)
    def gate(k):
        ok_r = np.isfinite(ratio[k]) & (ratio[k] <= learned.ACCEPT)
        ok_m = np.isfinite(mis[k]) & (mis[k] <= 1e-3)
        return ok_r, ok_m

    def fallback(rejected):
        """The rejected shots again with the 20-cell PML, in one batch: wall
        seconds and trace misfits."""
        if not rejected.size:
            return 0.0, []
        fb = [lib.Simulation(jobs[20][i]) for i in rejected]
        fbb = lib.Batch(fb)
        fbb.set_schedule(256, lib.SCHED_GRID, True)
        a = time.perf_counter()
        fbb.run(total)
        t = time.perf_counter() - a
        m = [misfit(s, sims["wide"][i]) for s, i in zip(fb, rejected)]
        fbb.close()
        for s in fb:
            s.close()
        return t, m

    ok_refl, ok_mis = gate("learned")
    accepted = ok_refl & ok_mis
    rejected = np.flatnonzero(~accepted)
    t_fb, fb_mis = fallback(rejected)
    final = mis["learned"].copy()
    final[rejected] = fb_mis
    t_learned = wall["learned"] + t_fb
    # the same gate and fallback with the plain 8-cell PML in the strip
    p_r, p_m = gate("plain8")
    p_rej = np.flatnonzero(~(p_r & p_m))
    t_pfb, _ = fallback(p_rej)

    def stats(x):
        x = np.asarray(x, np.float64)
        fin = x[np.isfinite(x)]
        return {"p50": float(np.percentile(fin, 50)) if fin.size else None,
                "p99": float(np.percentile(fin, 99)) if fin.size else None,
                "max": float(fin.max()) if fin.size else None, "non_finite": int((~np.isfinite(x)).sum())}

    out = {
        "metric": "shots/s batched vs plain at <=1e-3 L2 misfit",
        "value": round(n / t_learned, 3),
        "unit": "shots/s",
        "n_gpus": 1, "steps": total, "warmup": 0,
        "ms_per_step": round(t_learned * 1e3 / total, 4),
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (network: bf16 weights and activations, f32 accumulation)",
        "data": "synthetic",
        "config": {"workload": f"{n} held-out shots, layered half-spaces 500..3000 x 500..1500 cells, "
                               f"dx=2 m, learned 8-cell boundary + batching; judge: 40-cell PML",
                   "shots": n, "cells_per_step_strip8": sum(cells[8]),
                   "network": {"residual": net.residual, "damping": net.damping, "gain": net.gain,
                               "inner": net.inner, "profile": net.profile}},
        "value_basis": "all shots delivered at <= 1e-3 trace misfit: the learned batch plus the "
                       "fallback batch of the rejected shots",
        "plain_gpu_form_shots_per_s": round(n / t_plain, 3),
        "speedup_vs_plain_gpu_form": round(t_plain / t_learned, 3),
        "batched_shots_per_s": {k: round(n / wall[k], 3) for k in wall},
        "fallback_s": round(t_fb, 3),
        "batch_device_ms": {k: round(v, 1) for k, v in dev_ms.items()},
        "plain_shot_s": {"p50": round(float(np.percentile(per_shot, 50)), 4),
                         "p99": round(float(np.percentile(per_shot, 99)), 4)},
        "reflected_energy_vs_wide_pml": {k: stats(v) for k, v in ratio.items()},
        "trace_rel_l2_vs_wide_pml": {**{k: stats(v) for k, v in mis.items()}, "delivered": stats(final)},
        "learned_accepted": int(accepted.sum()),
        "plain8_same_gate": {"accepted": int(n - p_rej.size),
                             "shots_per_s": round(n / (wall["plain8"] + t_pfb), 3)},
        "rejected_reflected_energy": int((~ok_refl).sum()),
        "rejected_trace_misfit": int((~ok_mis).sum()),
        "judge": "FP32 40-cell PML run on the GPU; its own FP32-vs-FP64 error is stated separately "
                 "(relative L2 ~1e-5 or less), the FP64 oracle itself is not run on all 64 shots",
        "roofline": None, "cpu_baseline": None,
    }
    for b in batches.values():
        b.close()
    for v in sims.values():
        for s in v:
            s.close()
    net.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("roofline", "batched", "learned"), default="roofline",
                    help="roofline: the 8000 x 4000 model (default); batched: many shots; "
                         "learned: many shots with the learned 8-cell boundary against a 40-cell PML")
    ap.add_argument("--shots", type=int, default=64)
    ap.add_argument("--lb-gain", type=float, default=None,
                    help="learned workload: output gain (default: the saved network's)")
    ap.add_argument("--lb-damping", type=int, choices=(0, 1), default=None,
                    help="learned workload: 1 damping form, 0 residual form (default: the saved network's)")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--mx", type=int, default=8000)
    ap.add_argument("--mz", type=int, default=4000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sweep", type=int, choices=(1, 2), default=1,
                    help="roofline workload: steps per sweep (1, default: one step per launch "
                         "sequence; 2: the interior two steps per pass, bq_set_sweep)")
    ap.add_argument("--tiles", choices=("auto", "fused", "cells", "tune"), default="auto",
                    help="roofline workload: fused interior tiles or not (bq_set_tiles; auto: "
                         "fused where the grid has enough of them; tune: bq_autotune over "
                         "35 of the warm-up steps)")
    ap.add_argument("--materials", choices=("indexed", "cell"), default="indexed",
                    help="roofline workload: quantized material table (default) or one "
                         "material per cell (184 B per cell-step basis; no CPU baseline)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also time a plain 1 GiB device copy (k_copy_probe), the fused "
                         "tiles' access pattern without arithmetic (k_march_probe) and its "
                         "per-cell traffic without the march in 1 and in the march's number "
                         "of row bands (k_bands_probe)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.workload == "learned":
        args.steps = 3000 if args.steps is None else args.steps
        return run_learned(args)
    if args.workload == "batched":
        args.steps = 3000 if args.steps is None else args.steps
        args.warmup = 12 if args.warmup is None else args.warmup
        return run_batched(args)
    args.steps = 2000 if args.steps is None else args.steps
    args.warmup = 100 if args.warmup is None else args.warmup

    total = args.warmup + args.steps
    t0 = time.perf_counter()
    per_cell = args.materials == "cell"
    b_step, b_half, b_min = ((B_STEP_CELL, B_HALF_CELL, B_FUSED_MIN_CELL) if per_cell
                             else (B_STEP, B_HALF, B_FUSED_MIN))
    job = make_workload(total, args.mx, args.mz, per_cell)
    if not per_cell and job.table.shape[0] <= 65535:
        b_min = B_FUSED_MIN_IDX16
    nz, nx = hostsetup.grid_shape(job)
    print(f"[bench] model ready: {nx}x{nz} cells, JMNUM={job.table.shape[0]} "
          f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    sim = lib.Simulation(job)
    sim.set_sweep(args.sweep)
    tune = None
    warm = args.warmup
    if args.tiles == "tune":
        n_tune = len(lib.TUNE_TILES) * (TUNE_TRIAL + 1)
        if warm < n_tune:
            raise SystemExit(f"--tiles tune needs --warmup {n_tune} or more")
        tune = {c: round(v, 5) for c, v in sim.autotune(TUNE_TRIAL).items()}
        warm -= n_tune
        print(f"[bench] autotune ms/step: {tune}", file=sys.stderr, flush=True)
    else:
        sim.set_tiles(args.tiles)
    tiles = sim.tiles()
    print(f"[bench] uploaded; {sim.sweep()} step(s) per sweep, {tiles} tiles", file=sys.stderr, flush=True)
    sim.run(warm)
    print("[bench] warmup done", file=sys.stderr, flush=True)
    # HIP events around the launches of every 8th step of the timed region
    sim.enable_timing(TIMING_STRIDE)
    sim.sync()
    t1 = time.perf_counter()
    sim.run(args.steps, sync=False)
    sim.sync()
    t2 = time.perf_counter()
    stats = sim.kernel_stats()   # ms per launch and cells of each kind of launch
    ktimes, kcells, sweep = stats["ms"], stats["cells"], stats["steps_per_sweep"]
    elapsed = t2 - t1
    ms_step = elapsed * 1e3 / args.steps
    cells = nx * nz
    gcell = cells * args.steps / elapsed / 1e9
    u = sim.field("U")
    if not np.isfinite(u).all():
        raise SystemExit("non-finite field after run")
    if sweep == 2:
        # k_sweep2 advances the region's cells by two whole steps (2 x 104 B
        # each, 184 with per-cell materials) in one pass; it moves at least
        # 68 B (152) per region cell, once for both steps
        dom_name = "k_sweep2"
        dom_bytes = 2 * b_step * kcells["region"]
        min_bytes = b_min * kcells["region"]
        basis = (f"2 x {b_step} B per region cell (two steps per k_sweep2 launch); the other "
                 f"cells' launches are timed separately")
    else:
        # k_step advances the fused cells by a whole step (104 B each, 184 with
        # per-cell materials) and the other cells by the stress half-step (52 B, 88)
        dom_name = "k_step"
        dom_bytes = b_step * kcells["k_fused"] + b_half * kcells["k_stress"]
        min_bytes = b_min * kcells["k_fused"] + b_half * kcells["k_stress"]
        basis = f"{b_step} B per fused cell-step + {b_half} B per other stress cell, per k_step launch"
    # the first launch of a step: k_step, or k_rest_stress (the stress half-step
    # of every cell) on a grid without fused tiles
    dom_kernel = "k_rest_stress" if sweep != 2 and tiles == "cells" else dom_name
    if dom_kernel == "k_rest_stress":
        basis = f"{b_half} B per stress cell, per k_rest_stress launch (no fused tiles)"
    dom_s = ktimes[dom_name] * 1e-3
    achieved = dom_bytes / dom_s / 1e9
    traffic = None
    tr_path = os.path.join(ROOT, "profiles", "traffic.json")
    # the counted traffic of the 8000 x 4000 grid's launches with fused tiles
    if os.path.exists(tr_path) and not per_cell and (args.mx, args.mz) == (8000, 4000) and tiles == "fused":
        with open(tr_path) as f:
            traffic = json.load(f).get(dom_name)
    if traffic is not None:
        traffic = int(traffic)

    copy_gbs = march_gbs = bands_gbs = None
    if args.copy_probe:
        from biotfd import probes
        copy_gbs = probes.probe_copy(1 << 28, 20)
        march_gbs, _ = probes.probe_march(nx, nz, 20)
        # the march's per-cell traffic without the march, in row-major order and
        # over as many row bands as the march probe's rows of blocks
        march_bands = max(1, (nz - 8) // 132)
        bands_gbs = {str(nb): round(probes.probe_bands(nx, nz, nb, 20), 1) for nb in (1, march_bands)}

    cpu = cpu_all = None
    if not args.no_cpu_baseline and not per_cell:
        per_step = cpu_reference_baseline(job)
        if per_step and per_step > 0:
            cpu = {"value": round(cells / per_step / 1e9, 5), "unit": "Gcell-updates/s", "cores": 1,
                   "kind": "reference",
                   "sample": f"reference fd2d (flang -O2 -ffp-contract=off, serial) on the same "
                             f"{nx}x{nz} grid, 12 minus 2 steps, no output"}
        try:
            v, th = cpu_port_baseline(job)
            cpu_all = {"value": round(v, 5), "unit": "Gcell-updates/s", "cores": th, "kind": "port",
                       "sample": f"binary32 oracle (C restatement of the serial loop, OpenMP over "
                                 f"rows) on the same {nx}x{nz} grid, 3 steps after 1"}
        except Exception as exc:  # the baseline is reported, never required
            print(f"[bench] all-core baseline skipped: {exc}", file=sys.stderr, flush=True)

    out = {
        "metric": "Gcell-updates/s & %HBM roofline",
        "value": round(gcell, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"layered+correlated Biot half-space {args.mx}x{args.mz}, dx=2 m, "
                               f"TPML=20, free surface (stress imaging), "
                               f"{'one material per cell' if per_cell else 'materials quantized to 16 levels'}",
                   "grid_cells": cells, "nxc": nx, "nzc": nz, "jmnum": int(job.table.shape[0]),
                   "b_alg_bytes_per_cell_step": b_step, "receivers": len(job.receivers)},
        # whole step, all launches: B_alg x NXc x NZc per step over ms_per_step
        "hbm_roofline_frac_step": round(gcell * 1e9 * b_step / (HBM_PEAK_GBS * 1e9), 4),
        "steps_per_sweep": sweep,
        "tiles": tiles,
        "tile_rows": sim.tile_rows() if tiles == "fused" else None,
        **({"autotune_ms_per_step": tune} if tune else {}),
        "kernel_ms": {n: round(t, 5) for n, t in ktimes.items()},
        "kernel_cells": kcells,
        "roofline": {"bound": "hbm", "kernel": dom_kernel,
                     "basis": basis,
                     "algorithmic_bytes": int(dom_bytes),
                     "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        # the least the dominant launch can move (68 B per fused cell, 152
        # with per-cell materials, once per launch): its rate
        "k_step_min_bytes": {"bytes": int(min_bytes), "gbs": round(min_bytes / dom_s / 1e9, 1),
                             "frac": round(min_bytes / dom_s / 1e9 / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    if copy_gbs is not None:
        out["copy_probe_gbs"] = round(copy_gbs, 1)
        out["march_probe_gbs"] = round(march_gbs, 1)
        out["bands_probe_gbs"] = bands_gbs
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
