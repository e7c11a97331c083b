"""Batched shots (bq_batch_*) against the same shots run alone with bq_run,
bit for bit (needs a GPU).

The shots differ in grid size, PML width, surface treatment, bottom boundary,
source position and receivers; some have fused interior tiles (full, half and
masked quarter tiles), some only the per-cell and wavefront paths; one has a
PML too wide for the velocity launch's edge blocks, so its edge rows take the
kb_edges launch beside the others'. Every schedule - tile height 64 / 128 / 256,
one block per item or persistent, with and without the HIP graph - must give
each shot exactly the fields and traces of its own run.
"""
import numpy as np
import pytest

import jobs
from biotfd import lib

pytestmark = pytest.mark.gpu

FIELDS = ("U", "W", "QX", "QZ", "TXX", "TZZ", "TXZ", "P")
STEPS = 60


def _ktbo_job():
    job = jobs.tiny_layered(steps=STEPS, seed=21)
    job.ktbo = 4
    job.wc = 40.0
    return job


def _shot_jobs():
    return [
        jobs.wide_deep(steps=STEPS),                       # two 512-column fused tiles
        jobs.tiny_layered(steps=STEPS),
        jobs.tiny_edges(steps=STEPS),                      # source in the PML, receivers outside
        jobs.medium_layered(steps=STEPS),                  # wavefront runs
        jobs.tiny_layered_nosi(steps=STEPS, variant=3),    # explicit surface, AFDA options
        _ktbo_job(),                                       # bottom boundary key 4
        jobs.wide_deep(steps=STEPS, stress_imaging=False),
        jobs.narrow_tiles(steps=STEPS, mx=1000),           # half and masked quarter tiles
        jobs.tiny_layered(steps=STEPS, tpml=130),          # edge rows in kb_edges
    ]


def _compare(a, b, where):
    for f in FIELDS:
        x, y = a.field(f), b.field(f)
        bad = np.argwhere(~((x == y) | (np.isnan(x) & np.isnan(y))))
        assert bad.size == 0, f"{where}: {f} differs at {len(bad)} cells, first {bad[0]}"
    np.testing.assert_array_equal(a.traces(), b.traces(), err_msg=f"{where}: traces")


@pytest.mark.parametrize("rows,sched,graph", [
    (128, lib.SCHED_GRID, True), (64, lib.SCHED_PERSISTENT, True),
    (256, lib.SCHED_GRID, False), (128, lib.SCHED_PERSISTENT, False),
])
def test_batch_matches_shots_run_alone(rows, sched, graph):
    print(f"schedule rows={rows} sched={sched} graph={graph}", flush=True)
    js = _shot_jobs()
    batched = [lib.Simulation(j) for j in js]
    alone = [lib.Simulation(j) for j in js]
    # shot 0 starts 3 steps ahead: the shots of a batch need not be in step
    batched[0].run(3)
    b = lib.Batch(batched)
    b.set_schedule(rows, sched, graph)
    assert b.schedule() == (rows, sched, graph)
    for n in (1, 10, 7, 39):    # odd and even runs: graph replays and a remainder
        b.run(n)
    assert batched[0].step == 60 and batched[1].step == 57
    assert b.last_run_ms() > 0
    for s, a in zip(batched, alone):
        a.run(s.step)
    for k, (s, a) in enumerate(zip(batched, alone)):
        _compare(s, a, f"shot {k} ({js[k].name} {js[k].mx}x{js[k].mz})")
    # the shots carry on alone after the batch
    for s, a in zip(batched[1:], alone[1:]):
        s.run(3)
        a.run(3)
    _compare(batched[1], alone[1], "shot 1 after the batch")
    b.close()


def test_autotune_picks_a_candidate_and_keeps_results():
    js = _shot_jobs()[:4]
    batched = [lib.Simulation(j) for j in js]
    alone = [lib.Simulation(j) for j in js]
    b = lib.Batch(batched)
    times = b.autotune(trial_steps=2)
    print("autotune ms/step:", {k: round(v, 4) for k, v in times.items()}, flush=True)
    assert len(times) == 6 and all(v > 0 for v in times.values())
    rows, sched, graph = b.schedule()
    assert times[(rows, sched)] == min(times.values())
    b.run(STEPS - 12)
    for s, a in zip(batched, alone):
        a.run(STEPS)
    for k, (s, a) in enumerate(zip(batched, alone)):
        _compare(s, a, f"shot {k} after autotune")


def test_batch_rejects_steps_past_a_shots_end():
    js = _shot_jobs()[:2]
    sims = [lib.Simulation(j) for j in js]
    sims[1].run(STEPS - 5)
    b = lib.Batch(sims)
    with pytest.raises(lib.BiotError, match="beyond its MT2"):
        b.run(6)
    b.run(5)
    with pytest.raises(lib.BiotError):
        lib.Batch([sims[0], sims[0]])


def test_trace_misfit_on_device_matches_host():
    """bq_trace_misfit against the same sums formed on the host from the traces:
    two runs of one model with different PML widths and source frequencies
    (the same receivers)."""
    ja = jobs.medium_layered(steps=STEPS)
    jb = jobs.medium_layered(steps=STEPS, f0=55.0)
    jb.tpml = 16
    a, b = lib.Simulation(ja), lib.Simulation(jb)
    a.run(STEPS)
    b.run(STEPS)
    ta, tb = a.traces().astype(np.float64), b.traces().astype(np.float64)
    for first, n in ((1, STEPS), (11, 20)):
        tot, per = lib.trace_misfit(a, b, first, n, per_receiver=True)
        da, db = ta[first - 1:first - 1 + n], tb[first - 1:first - 1 + n]
        want_per = np.stack([((da - db) ** 2).sum(axis=(0, 2)), (db ** 2).sum(axis=(0, 2))], axis=1)
        np.testing.assert_allclose(per, want_per, rtol=1e-12, atol=0)
        np.testing.assert_allclose(tot, want_per.sum(axis=0), rtol=1e-12, atol=0)
        assert tot[0] > 0 and tot[1] > 0
    assert lib.trace_misfit(a, a)[0] == 0.0
    with pytest.raises(lib.BiotError, match="steps not run"):
        lib.trace_misfit(a, b, 1, STEPS + 1)


def test_batch_with_learned_boundary_matches_shots_run_alone():
    """Shots with a learned boundary attached (8-cell strip) in a batch with a
    plain shot: bit-identical to each shot run alone, the strip update
    included."""
    from biotfd import learned
    w = learned.init_weights(5)
    # small gains keep these untrained networks' runs finite over the steps
    nets = [learned.Network(w, np.full((3, 8), 0.8, np.float32), residual=False, gain=0.02),
            learned.Network(w, np.full((3, 8), 1.3, np.float32), residual=True, gain=0.05, inner=3),
            learned.Network(w, np.full((3, 8), 1.1, np.float32), residual=True, gain=1.0, damping=True)]
    js = [jobs.tiny_layered(steps=STEPS, tpml=8), jobs.medium_layered(steps=STEPS, tpml=8),
          jobs.tiny_layered(steps=STEPS, tpml=8), jobs.tiny_edges(steps=STEPS)]

    def make():
        sims = [lib.Simulation(j) for j in js]
        for k, s in enumerate(sims[:3]):
            rng = np.random.default_rng(40 + k)
            for f in FIELDS:
                s.set_field(f, (0.01 * rng.standard_normal((s.nz, s.nx))).astype(np.float32))
            nets[k].attach(s)
        return sims

    batched, alone = make(), make()
    b = lib.Batch(batched)
    b.set_schedule(128, lib.SCHED_GRID, True)
    b.run(1)
    b.run(8)     # four graph replays, one step without
    for a in alone:
        a.run(9)
    for k, (s, a) in enumerate(zip(batched, alone)):
        _compare(s, a, f"learned-boundary shot {k}")
    for s in batched:
        assert np.isfinite(s.field("U")).all() and np.abs(s.field("U")).max() > 0
    b.close()
