#!/usr/bin/env python3
"""States the binary32 error of the GPU time loop against the binary64 oracle
(relative L2 of the last snapshot of each field and of each receiver column
over all steps) on long runs, printing one JSON line per case:

* the roofline model (8000 x 4000 physical cells, layered x correlated field,
  quantized materials) on its full grid for --steps-full steps;
* the same model statistics on a 1000 x 500 grid for its full 8000 steps
  (--steps-long), so that the error growth over the workload's whole record
  is stated at a size the binary64 oracle finishes in minutes;
* optionally the 4000 x 2000 layered half-space on its full grid
  (--steps-layered, 5000 for its whole record);
* the learned-boundary workload's judge (40-cell PML, binary32 on the GPU)
  and the runs it judges, against the binary64 oracle's 40-cell PML run, on
  the smallest held-out shots (--judge-shots).

Needs a GPU; the oracle runs on the host cores (OMP_NUM_THREADS).

    python tests/fp64_statement.py [--steps-full N] [--steps-long N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [os.path.join(ROOT, "2d-vqsp-fd-sim_amd"), os.path.join(ROOT, "oracle"), ROOT]

import bench  # noqa: E402
import oracle  # noqa: E402
from biotfd import hostsetup, lib, models  # noqa: E402

FIELDS = ("U", "W", "QX", "QZ", "TXX", "TZZ", "TXZ", "P")
COLS = ("U", "-W", "QX", "-QZ", "PRES", "-TXX")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a))


def state(name, job):
    nz, nx = hostsetup.grid_shape(job)
    t0 = time.perf_counter()
    sim = lib.Simulation(job)
    sim.run(job.mt2)
    tg = time.perf_counter() - t0
    print(f"[fp64] {name}: GPU {job.mt2} steps in {tg:.1f} s", file=sys.stderr, flush=True)
    ref = oracle.OracleSim(job, "f64")
    t0 = time.perf_counter()
    left = job.mt2
    while left:
        n = min(left, 500)
        ref.run(n)
        left -= n
        print(f"[fp64] {name}: oracle at step {job.mt2 - left} ({time.perf_counter() - t0:.0f} s)",
              file=sys.stderr, flush=True)
    out = {"case": name, "grid": [nx, nz], "steps": job.mt2, "receivers": len(job.receivers),
           "jmnum": int(job.table.shape[0]),
           "fields_rel_l2": {f: rel_l2(sim.field(f), ref.field(f)) for f in FIELDS}}
    ta, tb = sim.traces(), ref.traces()
    out["traces_rel_l2"] = {c: rel_l2(ta[:, :, k], tb[:, :, k]) for k, c in enumerate(COLS)}
    out["oracle_s"] = round(time.perf_counter() - t0, 1)
    print(json.dumps(out), flush=True)
    sim.close()
    ref.close()


def judge(n_shots, steps):
    """The learned-boundary workload's judge against binary64: on the n
    smallest of its 64 held-out shots (seeds from 20000), the receiver traces
    (all six columns pooled, relative L2) of the 40-cell PML run on the GPU
    (the judge the workload uses), of the 8-cell strip with the trained
    network, of the plain 8-cell and of the plain 20-cell PML, each against
    the binary64 oracle's 40-cell PML run."""
    from biotfd import learned
    wide = models.shot_set(64, steps, seed=20_000, tpml=40)
    order = np.argsort([np.prod(hostsetup.grid_shape(j)) for j in wide])[:n_shots]
    net = learned.Network.load()
    for i in order:
        i = int(i)
        jobs = {t: models.shot_set(64, steps, seed=20_000, tpml=t)[i] for t in (40, 8, 20)}
        ref = oracle.OracleSim(jobs[40], "f64")
        t0 = time.perf_counter()
        ref.run(steps)
        tref = ref.traces()
        ref.close()
        judge = lib.Simulation(jobs[40])
        runs = {"gpu_fp32_pml40": judge, "learned_strip8": net.strip_grid(jobs[8], judge),
                "plain_pml8": lib.Simulation(jobs[8]), "plain_pml20": lib.Simulation(jobs[20])}
        net.attach(runs["learned_strip8"])
        out = {"case": f"learned-boundary judge, held-out shot {i}", "grid_pml40": list(hostsetup.grid_shape(jobs[40]))[::-1],
               "steps": steps, "network": {"damping": net.damping, "gain": net.gain}, "oracle_s": 0.0,
               "traces_rel_l2_vs_fp64_pml40": {}}
        for k, sim in runs.items():
            sim.run(steps)
            out["traces_rel_l2_vs_fp64_pml40"][k] = rel_l2(sim.traces(), tref)
            sim.close()
        out["oracle_s"] = round(time.perf_counter() - t0, 1)
        print(json.dumps(out), flush=True)
    net.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps-full", type=int, default=600)
    ap.add_argument("--steps-long", type=int, default=8000)
    ap.add_argument("--judge-shots", type=int, default=2)
    ap.add_argument("--steps-layered", type=int, default=0,
                    help="also the 4000 x 2000 layered half-space for this many steps (5000: its record)")
    args = ap.parse_args()
    if args.steps_layered:
        state("layered half-space, full 4000x2000 grid",
              models.layered_job(name="LAYER", mx=4000, mz=2000, h=2.0, steps=args.steps_layered,
                                 n_layers=10, seed=1, tpml=20, f0=25.0, n_rec=1000))
    if args.judge_shots:
        judge(args.judge_shots, 3000)
    if args.steps_long:
        state("roofline model statistics, 1000x500 grid", bench.make_workload(args.steps_long, 1000, 500))
    if args.steps_full:
        state("roofline model, full 8000x4000 grid", bench.make_workload(args.steps_full))


if __name__ == "__main__":
    main()
