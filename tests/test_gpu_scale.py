"""The HIP path at the workloads' own sizes (needs a GPU).

* The CPU-runnable workload (homogeneous Biot full space, 400 x 400 cells,
  1500 steps, 100 receivers) in full: every snapshot and receiver file the
  reference writes, by SHA-256 digest (tests/golden/cfg0_digests.npz, written
  by the reference itself).
* The layered half-space (4000 x 2000) and the roofline model (8000 x 4000,
  with quantized materials, with one material per cell, JMNUM = 32.0 M, and
  with two steps per sweep) on their full grids for a bounded number of
  steps, and the layered model on grids that take other partitions (fused
  blocks of 24 and 48 rows; no fused tiles at 4.3 M cells), bit for bit
  against the
  binary32 oracle on the host: from random initial fields, so that every cell
  of every row class - side and bottom PML strips, their memories, the
  surface rows, the bottom row and its history, every fused tile - is active
  from the first step, and from rest with the workload's own source.
* The binary32 error against the binary64 oracle, stated (printed) for the
  full CPU-runnable workload.
"""
import hashlib
import os
import sys
import tempfile
import time

import numpy as np
import pytest

import oracle
from biotfd import formats, hostsetup, lib, models, writers

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
sys.path.insert(0, os.path.dirname(HERE))   # bench.py: the roofline workload

pytestmark = pytest.mark.gpu

FIELDS = ("U", "W", "QX", "QZ", "TXX", "TZZ", "TXZ", "P")


def _sha(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def _compare(sim, ref, where):
    for f in FIELDS:
        a, b = sim.field(f), ref.field(f)
        bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
        assert bad.size == 0, (f"{where}: field {f} differs at {len(bad)} cells, first "
                               f"{bad[0]}: {a[tuple(bad[0])]} vs {b[tuple(bad[0])]}")


def test_cpu_runnable_workload_full_size_matches_reference_digests(tmp_path):
    import make_golden
    z = np.load(os.path.join(make_golden.HERE, "cfg0_digests.npz"), allow_pickle=False)
    want = dict(zip(z["files"].tolist(), z["sha256"].tolist()))
    job = make_golden.cfg0_job()
    nz, nx = hostsetup.grid_shape(job)
    print(f"cfg0 full: {nx}x{nz} cells, {job.mt2} steps, {len(job.receivers)} receivers", flush=True)
    sim = lib.Simulation(job)
    inside = writers.receivers_inside(job.receivers, 1 - job.tpml, job.mx + job.tpml,
                                      job.mzp + job.tpml)
    w = writers.SeismogramWriter(str(tmp_path), job.dt, len(job.receivers), key_tlq=True,
                                 key_tlp=True, inside=inside)
    done = 0
    for step in z["snap_steps"]:
        step = int(step)
        sim.run(step - done)
        w.write_chunk(done + 1, sim.traces(done + 1, step - done))
        done = step
        writers.write_snapshots(str(tmp_path), step, sim.field("U"), sim.field("W"),
                                sim.field("QX"), sim.field("QZ"))
    got = {f: _sha(os.path.join(tmp_path, f)) for f in want}
    bad = sorted(f for f in want if got[f] != want[f])
    # the REC_Q derivative columns of receivers outside the grid are
    # uninitialised in the reference (store_seismograms.F90:37-45); all 100
    # receivers of this workload lie inside it
    assert inside.all()
    assert not bad, f"{len(bad)} of {len(want)} files differ from the reference's, e.g. {bad[:5]}"


def _random_start(job, sim, ref, seed):
    """The same random binary32 fields in the GPU run and the oracle (whose
    previous filtration velocities QXP, QZP are the current ones at a start)."""
    rng = np.random.default_rng(seed)
    nz, nx = hostsetup.grid_shape(job)
    for f in FIELDS:
        a = rng.standard_normal((nz, nx), dtype=np.float32)
        sim.set_field(f, a)
        ref.set_field(f, a)
        if f in ("QX", "QZ"):
            ref.set_field(f + "P", a)


def _cfg1_job(steps, mx=4000, mz=2000):
    return models.layered_job(name="LAYER", mx=mx, mz=mz, h=2.0, steps=steps, n_layers=10,
                              seed=1, tpml=20, f0=25.0, n_rec=1000)


# grids whose default partition differs from the two workloads': fused blocks
# of 24 and 48 rows (with half tiles of 12 / 24 rows and quarter tiles of 6 /
# 12), and no fused tiles with 2-row runs just under the 4.5 M-cell limit
OTHER_GRIDS = {"layered 6000x3000": (6000, 3000, "fused"), "layered 5000x2500": (5000, 2500, "fused"),
               "layered 3000x1400": (3000, 1400, "cells")}


def _cfg2_job(steps, per_cell=False):
    import bench
    return bench.make_workload(steps, per_cell=per_cell)


@pytest.mark.parametrize("which,steps,start", [
    ("layered 4000x2000", 24, "random"), ("layered 4000x2000", 300, "rest"),
    ("layered 6000x3000", 6, "random"), ("layered 5000x2500", 6, "random"),
    ("layered 3000x1400", 24, "random"),
    ("roofline 8000x4000", 12, "random"), ("roofline 8000x4000", 120, "rest"),
    ("roofline 8000x4000 one material per cell", 6, "random"),
    ("roofline 8000x4000 two steps per sweep", 13, "random"),
])
def test_full_grid_bitexact_vs_oracle(which, steps, start):
    t0 = time.perf_counter()
    if which in OTHER_GRIDS:
        job = _cfg1_job(steps, *OTHER_GRIDS[which][:2])
    else:
        job = (_cfg1_job(steps) if which.startswith("layered")
               else _cfg2_job(steps, per_cell="per cell" in which))
    nz, nx = hostsetup.grid_shape(job)
    print(f"{which} ({start}): {nx}x{nz} cells, JMNUM={job.table.shape[0]}, {steps} steps, "
          f"{len(job.receivers)} receivers", flush=True)
    sim = lib.Simulation(job)
    if which in OTHER_GRIDS:
        assert sim.tiles() == OTHER_GRIDS[which][2]
    if "two steps" in which:
        sim.set_sweep(2)
        assert sim.sweep() == 2
    ref = oracle.OracleSim(job, "f32")
    if start == "random":
        _random_start(job, sim, ref, seed=7)
    np.testing.assert_array_equal(sim.sourtf, ref.sourtf)
    done = 0
    for chunk in (1, steps - 1):   # both field sets as the current one
        sim.run(chunk)
        t1 = time.perf_counter()
        ref.run(chunk)
        done += chunk
        print(f"  step {done}: oracle {time.perf_counter() - t1:.1f} s", flush=True)
        _compare(sim, ref, f"{which} step {done}")
    np.testing.assert_array_equal(sim.traces(), ref.traces())
    assert np.abs(sim.field("U")).max() > 0
    print(f"  {which} ({start}) bit-identical, {time.perf_counter() - t0:.1f} s", flush=True)
    sim.close()
    ref.close()


def _rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_fp32_error_vs_fp64_oracle_cpu_runnable_workload_full():
    """States the binary32 error of the GPU run against the binary64 oracle on
    the CPU-runnable workload in full (1500 steps): relative L2 of the last
    snapshot of each field and of each receiver column over all steps."""
    import make_golden
    job = make_golden.cfg0_job()
    sim = lib.Simulation(job)
    ref = oracle.OracleSim(job, "f64")
    sim.run(job.mt2)
    ref.run(job.mt2)
    out = {f: _rel_l2(sim.field(f), ref.field(f)) for f in FIELDS}
    ta, tb = sim.traces(), ref.traces()
    cols = ("U", "-W", "QX", "-QZ", "PRES", "-TXX")
    out.update({f"trace {c}": _rel_l2(ta[:, :, k], tb[:, :, k]) for k, c in enumerate(cols)})
    print("FP32 (GPU) vs FP64 oracle, CPU-runnable workload, 1500 steps, relative L2:", flush=True)
    for k, v in out.items():
        print(f"  {k:10s} {v:.3e}", flush=True)
    # a stated figure, not a tolerance; the bound only catches a broken run
    assert all(np.isfinite(v) and v < 1e-2 for v in out.values()), out
