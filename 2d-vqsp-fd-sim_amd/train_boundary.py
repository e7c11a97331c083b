#!/usr/bin/env python3
"""Trains the learned absorbing boundary on one MI355X and checks it.

1. Records bands from --models seeded layered models of 1000 x 1000 cells
   (dx = 2 m, random source position, Ricker 10..40 Hz) run with a 40-cell
   PML, in batches: groups of five consecutive levels every --every steps
   from --first on, over --steps steps.
2. Scales per side and field (RMS), then Adam on each side's network
   (--iters minibatches of 8 groups, cosine-decayed learning rate --lr),
   with a tenth of the groups held out for the validation loss.
3. Runs --check held-out models (seeds from 10000) several ways - 40-cell
   PML (the judge), 8-cell strip with the network at each of --gains (in the
   residual and, with --damping, the damping form), 8-cell plain PML - and
   prints the reflected energy of the others against the judge.

Writes the networks to --out (npz: weights, scales, loss history, check).
With --load, steps 1 and 2 are skipped and a saved network is checked.

    python 2d-vqsp-fd-sim_amd/train_boundary.py --out weights.npz
    python 2d-vqsp-fd-sim_amd/train_boundary.py --load weights.npz --damping --gains 1,0.5
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from biotfd import learned, lib  # noqa: E402


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def check(net, gains, n, steps, every, size=1000, forms=(False,)):
    """Reflected energy of the learned boundary at each gain (and form: damping
    False/True), of the plain 8-cell boundary and of the network's strip
    without the network against the 40-cell PML run on n held-out models."""
    out = []
    for k in range(n):
        seed = 10_000 + k
        wide = lib.Simulation(learned.training_job(seed, steps, 40, size, size))
        job8 = lambda: learned.training_job(seed, steps, 8, size, size)
        runs = {"plain8": lib.Simulation(job8())}
        if net.profile != "cpml8":
            runs[f"{net.profile} alone"] = net.strip_grid(job8(), wide)
        nets = []
        for damp in forms:
            for g in gains:
                s = net.strip_grid(job8(), wide)
                nets.append(net.with_gain(g, damping=damp))
                nets[-1].attach(s)
                runs[f"{'damping' if damp else 'learned'}@{g:g}"] = s
        acc = {m: np.zeros(2) for m in runs}
        done = 0
        for t in range(every, steps + 1, every):
            for s in [wide] + list(runs.values()):
                s.run(t - done)
            done = t
            for m, s in runs.items():
                acc[m] += learned.reflected_energy(s, wide)
        r = {m: float(v[0] / v[1]) for m, v in acc.items()}
        log(f"[lb] check model {seed}: " + ", ".join(f"{m} {v:.3e}" for m, v in r.items()))
        out.append(r)
        for s in [wide] + list(runs.values()):
            s.close()
        for x in nets:
            x.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--models", type=int, default=200)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--first", type=int, default=300)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--iters", type=int, default=2000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--check", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default="lb_strip8.npz")
    ap.add_argument("--direct", action="store_true",
                    help="train the direct form (the network is the strip's whole update) "
                         "instead of the residual form (a correction of the 8-cell CPML)")
    ap.add_argument("--inner", type=int, default=8,
                    help="strip cells nearest the interior that the network corrects (1..8)")
    ap.add_argument("--gains", default="1,0.5,0.25",
                    help="output gains to check on the held-out models")
    ap.add_argument("--damping", action="store_true",
                    help="also check the damping form of the residual network")
    ap.add_argument("--load", default=None, help="check this saved network only")
    ap.add_argument("--profile", choices=learned.PROFILES, default="cpml8",
                    help="the 8-cell strip's PML: the 8-cell CPML or the inner 8 nodes of the "
                         "40-cell profile")
    args = ap.parse_args()
    gains = [float(g) for g in args.gains.split(",")]
    forms = (False, True) if args.damping else (False,)
    if args.load:
        net = learned.Network.load(args.load)
        res = check(net, gains, args.check, args.steps, 500, forms=forms)
        print(json.dumps({"check": res}), flush=True)
        return
    t0 = time.perf_counter()
    per_model = len(range(args.first, args.steps - 1, args.every))
    tr = learned.Trainer(1000, 1002, max_groups=args.models * per_model, batch=args.batch,
                         residual=not args.direct, inner=args.inner)
    ng = learned.record(tr, args.models, args.steps, args.first, args.every, log=log,
                        profile=args.profile)
    t_rec = time.perf_counter() - t0
    scales = tr.compute_scales(ng)
    log(f"[lb] {ng} groups in {t_rec:.1f} s; scales {np.array2string(scales, precision=3)}")
    t1 = time.perf_counter()
    hist = learned.train(tr, learned.init_weights(args.seed), args.iters, args.lr, seed=args.seed, log=log)
    t_train = time.perf_counter() - t1
    net = learned.Network(tr.get_weights(), scales, residual=not args.direct, inner=args.inner,
                          profile=args.profile)
    res = check(net, gains, args.check, args.steps, 500, forms=forms) if args.check else []
    summary = {"groups": ng, "record_s": round(t_rec, 1), "train_s": round(t_train, 1),
               "iters": args.iters, "lr": args.lr, "residual": not args.direct,
               "inner": args.inner, "profile": args.profile,
               "final_train_loss": [h[-1] for h in hist["train"]],
               "final_val_loss": [h[-1] for h in hist["val"]],
               "zero_output_val_loss": hist["zero_val"],
               "check": res}
    net.save(args.out, train_loss=np.array(hist["train"]), val_loss=np.array(hist["val"]),
             info=json.dumps(summary))
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
