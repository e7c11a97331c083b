"""HIP path vs the CPU oracle, bit for bit, through the C ABI (needs a GPU)."""
import os
import sys

import numpy as np
import pytest

import jobs
import oracle
from biotfd import lib

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))

pytestmark = pytest.mark.gpu

FIELDS = ("U", "W", "QX", "QZ", "TXX", "TZZ", "TXZ", "P")


def _compare(sim, ref, where):
    for f in FIELDS:
        a, b = sim.field(f), ref.field(f)
        bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
        assert bad.size == 0, f"{where}: field {f} differs at {len(bad)} cells, first {bad[0]}: {a[tuple(bad[0])]} vs {b[tuple(bad[0])]}"


CASES = {**jobs.ALL, **jobs.GPU_EXTRA}


@pytest.mark.parametrize("name", sorted(CASES))
def test_bitexact_vs_oracle(name):
    """Each case twice, beside one oracle run: with the interior in fused tiles
    and with every cell on the runs and the per-cell path (bq_set_tiles)."""
    job = CASES[name]()
    print(f"case {name}: grid {job.mx}x{job.mz} T={job.tpml} steps={job.mt2}", flush=True)
    ref = oracle.OracleSim(job, "f32")
    sims = {}
    for mode in ("fused", "cells"):
        sims[mode] = lib.Simulation(job)
        sims[mode].set_tiles(mode)
    assert sims["cells"].tiles() == "cells"
    np.testing.assert_array_equal(sims["fused"].sourtf, ref.sourtf)
    done = 0
    for chunk in (1, 9, job.mt2 - 10):
        ref.run(chunk)
        done += chunk
        for mode, sim in sims.items():
            sim.run(chunk)
            _compare(sim, ref, f"{name} ({mode} tiles) step {done}")
    for sim in sims.values():
        np.testing.assert_array_equal(sim.traces(), ref.traces())
        assert np.abs(sim.field("U")).max() > 0


def test_pml_coefficients_match_oracle():
    job = jobs.tiny_layered(steps=2)
    ref = oracle.OracleSim(job, "f32")
    sim = lib.Simulation(job)
    for k in range(24):
        np.testing.assert_array_equal(sim.pml(k), ref.pml(k), err_msg=f"PML array {k}")


def test_bottom_nr_and_upload_pml():
    """KTBO != 0 path: explicit bottom coefficients, PML arrays uploaded from the host."""
    job = jobs.tiny_layered(steps=80)
    job.ktbo = 2
    job.wc = 40.0   # with KTBO != 0 the bottom damping is off and WC > 0 keeps ODC finite
    abw = np.array([0, 0, 0.9, 1.1, 0, 0, -1.1, 0.1], np.float32)
    abh = np.array([0.05, 0, 0.8, 1.2, 0, 0, -1.2, 0.2], np.float32)
    ref = oracle.OracleSim(job, "f32", abw=abw, abh=abh)
    arrs = [ref.pml(k) for k in range(24)]
    sim = lib.Simulation(job, abw=abw, abh=abh, pml_arrays=arrs)
    ref.run(job.mt2)
    sim.run(job.mt2)
    _compare(sim, ref, "ktbo")
    assert np.isfinite(sim.field("U")).all()


@pytest.mark.parametrize("key", range(10))
def test_bottom_nr_coefficients_match_oracle(key):
    """bq_build_bottom_nr (from the uploaded model) vs the oracle's PRE_NR_COEF1."""
    job = jobs.tiny_layered(steps=2)
    job.ktbo = key
    job.wc = 40.0
    job.omg = 6.283185307179586 * 30.0
    ref = oracle.OracleSim(job, "f32")
    sim = lib.Simulation(job)
    np.testing.assert_array_equal(sim.abw, ref.abw)
    np.testing.assert_array_equal(sim.abh, ref.abh)


def _golden_names():
    import make_golden
    return make_golden.names()


@pytest.mark.parametrize("name", _golden_names())
def test_bitexact_vs_reference_snapshots(name):
    """HIP path vs the snapshots the reference itself wrote (tests/golden)."""
    import make_golden
    job, z = make_golden.load(name)
    print(f"golden case {name}", flush=True)
    sim = lib.Simulation(job)   # KTBO != 0: the library computes PRE_NR_COEF1's coefficients
    done = 0
    for k, step in enumerate(z["snap_steps"]):
        sim.run(int(step) - done)
        done = int(step)
        for f, fname in enumerate(("U", "W", "QX", "QZ")):
            got, ref = sim.field(fname), z["snaps"][k, f]
            bad = np.argwhere(got != ref)
            assert bad.size == 0, f"{name} step {step} {fname}: {len(bad)} cells differ"


@pytest.mark.parametrize("every", [1, 3])
@pytest.mark.parametrize("tpml", [8, 130])
def test_kernel_timing_leaves_results_unchanged(every, tpml):
    """Timed runs (events around every n-th step) give the untimed fields and
    positive per-launch durations: k_edges is 0 where the velocity launch takes
    the edge rows (TPML 8) and positive where it runs on its own (TPML 130,
    too wide for the velocity launch's edge blocks)."""
    job = jobs.tiny_layered(tpml=tpml)
    print(f"timing every {every}, TPML {tpml}", flush=True)
    a, b = lib.Simulation(job), lib.Simulation(job)
    a.enable_timing(every)
    a.run(10)
    b.run(10)
    t = a.kernel_times()
    assert set(t) == {"k_step", "k_velocity", "k_edges"}
    assert t["k_step"] > 0 and t["k_velocity"] > 0, t
    assert (t["k_edges"] > 0) == (tpml > 124), t
    for f in FIELDS:
        np.testing.assert_array_equal(a.field(f), b.field(f))
    a.close()
    b.close()


def _second_wide():
    job = jobs.wide_deep(steps=70)
    job.variant = 4
    return job


AUTOTUNE_CASES = {"wide": lambda: jobs.wide_deep(steps=70),
                  "wide explicit surface": lambda: jobs.wide_deep(steps=70, stress_imaging=False),
                  "wide second-order build": _second_wide,
                  "narrow, masked quarter column": lambda: jobs.narrow_tiles(steps=70, mx=1000)}


@pytest.mark.parametrize("case", list(AUTOTUNE_CASES))
def test_autotune_switches_partitions_bit_exact(case):
    """bq_autotune runs real steps with fused blocks of 128, 64, 48, 32, 24 and
    16 rows and without fused tiles, switching between them from step to
    step, and keeps one: the run stays bit-identical to the oracle."""
    job = AUTOTUNE_CASES[case]()
    print(f"autotune {case}: grid {job.mx}x{job.mz}", flush=True)
    ref = oracle.OracleSim(job, "f32")
    sim = lib.Simulation(job)
    ms = sim.autotune(2)
    assert list(ms) == list(lib.TUNE_TILES)
    assert all(v > 0 for v in ms.values()), ms
    assert sim.step == 21
    best = min(ms, key=ms.get)
    assert sim.tiles() == ("cells" if best == "cells" else "fused"), (ms, sim.tiles())
    fused = {k: v for k, v in ms.items() if k != "cells"}
    assert sim.tile_rows() == int(min(fused, key=fused.get)[5:]), (ms, sim.tile_rows())
    ref.run(21)
    _compare(sim, ref, "after the autotuner")
    sim.run(job.mt2 - 21)
    ref.run(job.mt2 - 21)
    _compare(sim, ref, "end of run")
    np.testing.assert_array_equal(sim.traces(), ref.traces())


def test_tiles_modes_and_errors():
    """BQ_TILES_AUTO leaves a small grid without fused tiles, BQ_TILES_FUSED and
    BQ_TILES_CELLS force the partition; bad modes and autotuner calls that
    would run past MT2 are rejected."""
    job = jobs.wide_deep(steps=20)
    sim = lib.Simulation(job)
    assert sim.tiles() == "cells"
    assert sim.tile_rows() in (128, 64, 48, 32, 24, 16)
    sim.set_tiles("fused")
    assert sim.tiles() == "fused"
    sim.set_tiles("cells")
    assert sim.tiles() == "cells"
    sim.set_tiles("auto")
    assert sim.tiles() == "cells"
    assert sim.lib.bq_set_tiles(sim.h, 3) == -2
    assert sim.lib.bq_autotune(sim.h, 0, None) == -2
    assert sim.lib.bq_autotune(sim.h, 2, None) == -7   # 21 steps > MT2 = 20
    sim.close()


def test_set_receivers_replaced_before_first_step_rejected_after():
    """A second bq_set_receivers before the first step replaces the receivers
    (and frees the previous buffers); once a step has run it is rejected."""
    job = jobs.tiny_layered(steps=30)
    sim = lib.Simulation(job)
    rec = np.asarray(job.receivers, np.int32).reshape(-1, 4)[:3]
    cols = [np.ascontiguousarray(rec[:, k]) for k in range(4)]
    sim._chk(sim.lib.bq_set_receivers(sim.h, 3, *[lib._ptr(c) for c in cols]))
    sim.mr = 3
    sim.run(job.mt2)
    job3 = jobs.tiny_layered(steps=30)
    job3.receivers = [tuple(r) for r in rec.tolist()]
    ref = oracle.OracleSim(job3, "f32")
    ref.run(job3.mt2)
    np.testing.assert_array_equal(sim.traces(), ref.traces())
    rc = sim.lib.bq_set_receivers(sim.h, 3, *[lib._ptr(c) for c in cols])
    assert rc == -6
    assert b"after the first step" in sim.lib.bq_last_error(sim.h)
    np.testing.assert_array_equal(sim.traces(), ref.traces())
