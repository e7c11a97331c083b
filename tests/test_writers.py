"""Output writers (biotfd.writers) against the reference's own files.

The reference binary built from its sources (oracle/_ref/fd2d) writes the
receiver files and snapshots of a small job; the writers, fed with the oracle's
samples and fields (bit-identical to the reference's, see test_oracle_golden),
must reproduce them byte for byte. Skipped where the reference binary is absent.
"""
import filecmp
import os
import subprocess

import numpy as np
import pytest

import jobs
import oracle
from biotfd import formats, hostsetup, writers

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "fd2d")


@pytest.mark.parametrize("x,s", [
    (0.0, " 0.000000E+00"), (1.0, " 0.100000E+01"), (-2.5e-7, "-0.250000E-06"),
    (123456.7, " 0.123457E+06"), (9.9999996e-1, " 0.100000E+01"), (1e-40, " 0.999995E-40"),
])
def test_fortran_e_edit(x, s):
    assert writers.fortran_e(x) == s


@pytest.mark.skipif(not os.path.exists(REF), reason="reference binary not built")
@pytest.mark.parametrize("name", ["edges", "layered"])
def test_writers_match_reference_files(tmp_path, name):
    job = jobs.ALL[name]()
    job.key_tlv = job.key_tlq = job.key_tlp = job.key_snv = True
    job.ipas2 = 50
    ref_dir, our_dir = tmp_path / "ref", tmp_path / "ours"
    ref_dir.mkdir()
    our_dir.mkdir()
    formats.write_job(job, str(ref_dir))
    subprocess.run([REF], cwd=ref_dir, check=True, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL, timeout=300)
    sim = oracle.OracleSim(job, "f32")
    nz, nx = hostsetup.grid_shape(job)
    inside = writers.receivers_inside(job.receivers, 1 - job.tpml, job.mx + job.tpml, job.mzp + job.tpml)
    w = writers.SeismogramWriter(str(our_dir), job.dt, len(job.receivers), key_tlq=True, key_tlp=True)
    for step in range(1, job.mt2 + 1):
        sim.run(1)
        w.write(step, sim.traces(step, 1)[0], inside)
        if step % job.ipas2 == 0:
            writers.write_snapshots(str(our_dir), step, sim.field("U"), sim.field("W"),
                                    sim.field("QX"), sim.field("QZ"))
    names = sorted(f for f in os.listdir(ref_dir) if f.startswith(("REC_", "SNAP.")))
    assert len(names) == 3 * len(job.receivers) + 2 * (job.mt2 // job.ipas2)
    for f in names:
        a, b = ref_dir / f, our_dir / f
        if f.startswith("REC_Q_") and not inside[int(f[6:-4]) - 1].all():
            # the reference's derivative columns are uninitialised there
            la = [l[:42] for l in a.read_text().splitlines()]
            lb = [l[:42] for l in b.read_text().splitlines()]
            assert la == lb, f
        else:
            assert filecmp.cmp(a, b, shallow=False), f
