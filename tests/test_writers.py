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
    (-0.0, "-0.000000E+00"), (3.4028235e38, " 0.340282E+39"), (1.4e-45, " 0.140130E-44"),
    # the spellings the reference's runtime writes for a run that blew up
    # (test_writers_match_reference_on_blowup pins them against its files)
    (float("nan"), "          NaN"), (float("inf"), "          Inf"), (float("-inf"), "         -Inf"),
])
def test_fortran_e_edit(x, s):
    assert writers.fortran_e(x) == s


def _reference_and_ours(tmp_path, job, chunks):
    """Runs the reference on job and writes our files from the oracle's samples
    (bit-identical to the reference's) in chunks of the given sizes."""
    ref_dir, our_dir = tmp_path / "ref", tmp_path / "ours"
    ref_dir.mkdir()
    our_dir.mkdir()
    formats.write_job(job, str(ref_dir))
    subprocess.run([REF], cwd=ref_dir, check=True, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL, timeout=300)
    sim = oracle.OracleSim(job, "f32")
    inside = writers.receivers_inside(job.receivers, 1 - job.tpml, job.mx + job.tpml, job.mzp + job.tpml)
    w = writers.SeismogramWriter(str(our_dir), job.dt, len(job.receivers), key_tlq=job.key_tlq,
                                 key_tlp=job.key_tlp, inside=inside)
    step, k = 1, 0
    while step <= job.mt2:
        n = min(chunks[k % len(chunks)], job.mt2 - step + 1)
        if job.key_snv:   # a chunk ends at every snapshot step
            n = min(n, ((step + job.ipas2 - 1) // job.ipas2) * job.ipas2 - step + 1)
        sim.run(n)
        w.write_chunk(step, sim.traces(step, n))
        step += n
        k += 1
        if job.key_snv and (step - 1) % job.ipas2 == 0:
            writers.write_snapshots(str(our_dir), step - 1, sim.field("U"), sim.field("W"),
                                    sim.field("QX"), sim.field("QZ"))
    return ref_dir, our_dir, inside


@pytest.mark.skipif(not os.path.exists(REF), reason="reference binary not built")
@pytest.mark.parametrize("name,chunks", [("edges", [1]), ("layered", [7, 1, 50]), ("edges", [256])])
def test_writers_match_reference_files(tmp_path, name, chunks):
    job = jobs.ALL[name]()
    job.key_tlv = job.key_tlq = job.key_tlp = job.key_snv = True
    job.ipas2 = 50
    ref_dir, our_dir, inside = _reference_and_ours(tmp_path, job, chunks)
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


@pytest.mark.skipif(not os.path.exists(REF), reason="reference binary not built")
def test_writers_match_reference_on_blowup(tmp_path):
    """A run at three times the stable step blows up to Inf and NaN; the
    receiver files must still match the reference's byte for byte, which pins
    how the reference's runtime spells non-finite values."""
    job = jobs.tiny_homog(steps=400)
    job.dt = float(np.float32(job.dt * 3.0))
    job.mt = float(np.float32((400 - 0.5) * np.float32(job.dt)))
    job.key_tlv = job.key_tlq = job.key_tlp = True
    assert job.mt2 == 400
    ref_dir, our_dir, inside = _reference_and_ours(tmp_path, job, [64])
    assert inside.all()
    text = ""
    for f in sorted(f for f in os.listdir(ref_dir) if f.startswith("REC_")):
        assert filecmp.cmp(ref_dir / f, our_dir / f, shallow=False), f
        text += (ref_dir / f).read_text()
    assert "NaN" in text and " Inf" in text and "-Inf" in text
