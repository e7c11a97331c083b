"""The CPU oracle against the reference's own outputs (no GPU needed).

The fixtures hold snapshots written by the reference compiled from its sources
(tests/golden/make_golden.py). The binary32 oracle must reproduce them bit for
bit; the reference built with flang's default FMA contraction must agree within
the 1e-5 relative L2 parity bar; the FP64 oracle states the binary32 error.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402
import oracle  # noqa: E402

NAMES = make_golden.names()
FIELDS = ("U", "W", "QX", "QZ")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def test_fixtures_present():
    assert {"layered", "homog", "edges", "layered_ktbo2", "homog_fluid_source"} <= set(NAMES)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_bitexact_vs_reference(name):
    job, z = make_golden.load(name)
    sim = oracle.OracleSim(job, "f32")
    done = 0
    for k, step in enumerate(z["snap_steps"]):
        sim.run(int(step) - done)
        done = int(step)
        for f, fname in enumerate(FIELDS):
            ref = z["snaps"][k, f]
            got = sim.field(fname)
            bad = np.argwhere(~((got == ref) | (np.isnan(got) & np.isnan(ref))))
            assert bad.size == 0, (f"{name} step {step} {fname}: {len(bad)} cells differ, first "
                                   f"{bad[0]} {got[tuple(bad[0])]} vs {ref[tuple(bad[0])]}")
    assert np.isfinite(z["snaps"]).all() and np.abs(z["snaps"][-1]).max() > 0


@pytest.mark.parametrize("name", NAMES)
def test_reference_fma_build_within_parity_bar(name):
    """flang's default build contracts into FMA; its fields stay within 1e-5 rel L2."""
    job, z = make_golden.load(name)
    for f in range(4):
        assert rel_l2(z["snap_fma_last"][f], z["snaps"][-1, f]) < 1e-5


@pytest.mark.parametrize("name", NAMES)
def test_traces_match_reference_ascii(name):
    """Receiver samples vs the REC_*.asc files (E13.6: 6 significant digits)."""
    job, z = make_golden.load(name)
    sim = oracle.OracleSim(job, "f32")
    sim.run(job.mt2)
    tr = sim.traces().astype(np.float64)  # [step][mr][6]
    dt = np.float32(job.dt)
    for j in range(sim.mr):
        uw, q, st = z["rec_uw"][j], z["rec_q"][j], z["rec_stress"][j]
        t = (np.arange(1, job.mt2 + 1, dtype=np.float32) * dt).astype(np.float64)
        np.testing.assert_allclose(uw[:, 0], t, rtol=5.0001e-6)
        for col, (ref, k) in enumerate(((uw[:, 1], 0), (uw[:, 2], 1), (q[:, 1], 2), (q[:, 2], 3),
                                         (st[:, 1], 4), (st[:, 2], 5))):
            got = tr[:, j, k]
            # E13.6 prints 0.dddddd x 10^e: half a unit of the 6th digit is at
            # most 5e-6 of the printed value
            tol = 5.0001e-6 * np.abs(ref) + 1e-44
            assert np.all(np.abs(got - ref) <= tol), (name, j, col)


@pytest.mark.parametrize("name", ["layered", "homog", "edges"])
def test_fp32_error_vs_fp64_oracle(name):
    """The binary32 run (bit-identical to the reference) against the FP64 oracle."""
    job, z = make_golden.load(name)
    s64 = oracle.OracleSim(job, "f64")
    s64.run(job.mt2)
    for f, fname in enumerate(FIELDS):
        err = rel_l2(z["snaps"][-1, f], s64.field(fname))
        assert err < 1e-4, (fname, err)


def test_oracle_full_cpu_runnable_workload_matches_reference_digests(tmp_path):
    """The binary32 oracle on the CPU-runnable workload at its full size (400 x
    400 cells, 1500 steps, 100 receivers) writes every snapshot and receiver
    file byte for byte as the reference does (digests in cfg0_digests.npz)."""
    import hashlib
    from biotfd import hostsetup, writers
    z = np.load(os.path.join(HERE, "golden", "cfg0_digests.npz"), allow_pickle=False)
    want = dict(zip(z["files"].tolist(), z["sha256"].tolist()))
    job = make_golden.cfg0_job()
    sim = oracle.OracleSim(job, "f32")
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
    assert len(want) == 3 * len(job.receivers) + 2 * len(z["snap_steps"])
    for f, h in want.items():
        with open(tmp_path / f, "rb") as fh:
            assert hashlib.sha256(fh.read()).hexdigest() == h, f
