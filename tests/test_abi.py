"""C ABI checks that need no GPU: symbols, host-side helpers."""
import os
import re

import numpy as np
import pytest

import jobs
import oracle
from biotfd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header="biotfd.h"):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bq_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 19
    so = lib.load()
    for n in names:
        assert hasattr(so, n), n
    assert set(lib.EXPORTS) <= set(names)


@pytest.mark.parametrize("name", sorted(jobs.ALL))
@pytest.mark.parametrize("key_sour,qr", [(0, 0.0), (0, 0.25), (1, 0.0), (2, 0.4), (3, 0.0)])
def test_source_spline_matches_oracle(name, key_sour, qr):
    """OPEN_SR_FILE restated in the library's host code vs the oracle, bit for bit."""
    job = jobs.ALL[name]()
    job.key_sour, job.qr = key_sour, qr
    a = lib.source_from_stf(job)
    b = oracle.source_from_stf(job, "f32")
    np.testing.assert_array_equal(a, b)
    assert np.abs(a).max() > 0


def test_create_rejects_bad_descriptions():
    import ctypes as C
    so = lib.load()
    h = C.c_void_p()
    bad = [
        lib.Desc(4, 20, 8, 1, 2, 2, 0, 10, 1, 0, 1.0, 1e-4),     # MX too small
        lib.Desc(40, 30, 0, 1, 2, 2, 0, 10, 1, 0, 1.0, 1e-4),    # TPML 0
        lib.Desc(40, 30, 8, 0, 2, 2, 0, 10, 1, 8, 1.0, 1e-4),    # unknown variant bit
        lib.Desc(40, 30, 8, 1, 2, 29, 0, 10, 1, 0, 1.0, 1e-4),   # source row past MZ'-2
        lib.Desc(40, 30, 8, 1, 60, 5, 0, 10, 1, 0, 1.0, 1e-4),   # source column outside
    ]
    words = ["mx >= 8", "tpml", "variant", "source cell", "source cell"]
    for d, w in zip(bad, words):
        assert so.bq_create(C.byref(d), C.byref(h)) < 0
        assert h.value is None
        msg = so.bq_last_error(None).decode()
        assert msg.startswith("bq_create: ") and w in msg, msg


@pytest.mark.parametrize("key", range(10))
def test_nonref_matches_oracle(key):
    """NONREF restated in the library's host code vs the oracle, bit for bit,
    over a spread of speeds, cell sizes and steps (the oracle's NONREF is
    pinned by the reference's KTBO fixtures)."""
    rng = np.random.default_rng(100 + key)
    for _ in range(50):
        h = np.float32(rng.uniform(0.5, 5.0))
        vp = np.float32(rng.uniform(1500.0, 5000.0))
        vs = np.float32(vp * rng.uniform(0.3, 0.7))
        dt = np.float32(0.4 * h / vp)
        prm = np.array([rng.uniform(5.0, 80.0), rng.uniform(0.2, 0.6), 0.19757, 0.70511], np.float32)
        for vel in (vp, vs):
            a = lib.nonref(key, h, dt, vel, vs, vp, prm)
            b = oracle.nonref(key, h, dt, vel, vs, vp, prm, "f32")
            np.testing.assert_array_equal(a, b)
            assert np.isfinite(a).all()


def test_probes_library_is_separate():
    """The bandwidth probes live in their own library, not in the time-loop ABI."""
    from biotfd import probes
    src = open(os.path.join(ROOT, "include", "biotfd_probes.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(bq_[a-z_0-9]+)\s*\(", src)))
    assert names == sorted(probes.EXPORTS)
    so = probes.load()
    for n in names:
        assert hasattr(so, n), n
    assert not set(names) & set(declared_functions())
    assert not any(hasattr(lib.load(), n) for n in names)


def test_learned_boundary_entry_points_are_exported():
    """The learned-boundary extension (include/biotfd_learned.h) is part of the
    time-loop library and bound by biotfd.learned."""
    from biotfd import learned
    names = declared_functions("biotfd_learned.h")
    assert sorted(names) == sorted(learned.EXPORTS)
    so = lib.load()
    for n in names:
        assert hasattr(so, n), n
