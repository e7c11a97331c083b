"""Two steps per sweep (needs a GPU): bq_run with bq_set_sweep(h, 2) gives
fields and receiver traces bit-identical to single steps and to the binary32
oracle - from rest with the workload's source and from random fields (every
cell active), in chunks of odd and even lengths (a sweep never straddles two
runs), with the source inside the region on a tile's last column, with the
explicit free surface and the second-order build; a receiver inside the
region falls back to single steps."""
import numpy as np
import pytest

import jobs
import oracle
from biotfd import hostsetup, lib, models

pytestmark = pytest.mark.gpu

FIELDS = ("U", "W", "QX", "QZ", "TXX", "TZZ", "TXZ", "P")
CHUNKS = (1, 2, 7, 10, 3)


def _compare(a, b, where):
    for f in FIELDS:
        x, y = a.field(f), b.field(f)
        bad = np.argwhere(x != y)
        assert bad.size == 0, (f"{where}: {f} differs at {len(bad)} cells, first {bad[0]}: "
                               f"{x[tuple(bad[0])]} vs {y[tuple(bad[0])]}")


def _random(job, sims, seed):
    rng = np.random.default_rng(seed)
    nz, nx = hostsetup.grid_shape(job)
    for f in FIELDS:
        a = rng.standard_normal((nz, nx), dtype=np.float32)
        for s in sims:
            s.set_field(f, a)


def _big_layered(steps):
    return models.layered_job(name="BIG", mx=1300, mz=600, h=2.0, steps=steps, n_layers=5, seed=21,
                              tpml=20, f0=25.0, n_rec=300, field_std=0.05, corr_len=20.0, levels=4)


def _wide_second(steps):
    job = jobs.wide_deep(steps=steps)
    job.variant = 4
    return job


CASES = {
    "wide, source on a tile's last column": lambda: jobs.wide_deep(steps=40),
    "wide, explicit surface": lambda: jobs.wide_deep(steps=40, stress_imaging=False),
    "wide, second-order build": lambda: _wide_second(40),
    "1300x600 TPML 20": lambda: _big_layered(40),
}


@pytest.mark.parametrize("start", ["rest", "random"])
@pytest.mark.parametrize("case", list(CASES))
def test_two_steps_per_sweep_match_single_steps(case, start):
    job = CASES[case]()
    one, two = lib.Simulation(job), lib.Simulation(job)
    assert one.sweep() == 1    # the default
    two.set_sweep(2)
    assert two.sweep() == 2, "this grid should have a two-step region"
    if start == "random":
        _random(job, (one, two), seed=3)
    done = 0
    for n in CHUNKS:
        if done + n > job.mt2:
            break
        one.run(n)
        two.run(n)
        done += n
        _compare(two, one, f"{case} ({start}) step {done}")
    np.testing.assert_array_equal(two.traces(), one.traces())
    assert np.abs(two.field("U")).max() > 0
    one.close()
    two.close()


def test_two_steps_per_sweep_match_the_oracle():
    job = jobs.wide_deep(steps=30)
    sim = lib.Simulation(job)
    sim.set_sweep(2)
    assert sim.sweep() == 2
    ref = oracle.OracleSim(job, "f32")
    rng = np.random.default_rng(9)
    nz, nx = hostsetup.grid_shape(job)
    for f in FIELDS:
        a = rng.standard_normal((nz, nx), dtype=np.float32)
        sim.set_field(f, a)
        ref.set_field(f, a)
        if f in ("QX", "QZ"):
            ref.set_field(f + "P", a)
    for n in (4, 26):
        sim.run(n)
        ref.run(n)
        for f in FIELDS:
            np.testing.assert_array_equal(sim.field(f), ref.field(f), err_msg=f)
    np.testing.assert_array_equal(sim.traces(), ref.traces())
    sim.close()
    ref.close()


def test_receiver_inside_the_region_keeps_single_steps():
    job = jobs.wide_deep(steps=10)
    job.receivers = list(job.receivers) + [(500, 80, 500, 80)]
    sim = lib.Simulation(job)
    sim.set_sweep(2)
    assert sim.sweep() == 1
    sim.close()
    # a receiver on the region's rim (2 cells from its edge) keeps two steps
    job = jobs.wide_deep(steps=10)
    job.receivers = list(job.receivers) + [(500, 10, 500, 10)]
    sim = lib.Simulation(job)
    sim.set_sweep(2)
    assert sim.sweep() == 2
    sim.close()


def test_autotune_runs_single_steps_and_keeps_the_sweep_setting():
    """bq_autotune's trials are single steps whatever bq_set_sweep allows; the
    two-step setting holds again after it, and the run stays bit-identical to
    one that never left single steps."""
    job = jobs.wide_deep(steps=50)
    one, two = lib.Simulation(job), lib.Simulation(job)
    two.set_sweep(2)
    two.autotune(1)                 # 7 candidates x 2 steps
    assert two.sweep() == 2
    one.run(14)
    _compare(two, one, "after the autotuner")
    one.run(job.mt2 - 14)
    two.run(job.mt2 - 14)
    _compare(two, one, "end of run")
    np.testing.assert_array_equal(two.traces(), one.traces())
    one.close()
    two.close()
