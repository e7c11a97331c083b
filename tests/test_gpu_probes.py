"""Bandwidth probes of the library (diagnostics; need a GPU)."""
import pytest

from biotfd import lib

pytestmark = pytest.mark.gpu


def test_copy_probe_reports_bandwidth():
    assert lib.probe_copy(1 << 22, 3) > 0


def test_march_probe_reports_bandwidth():
    assert lib.probe_march(1100, 300, 3) > 0


def test_march_probe_rejects_small_grids():
    with pytest.raises(lib.BiotError):
        lib.probe_march(100, 300, 1)
