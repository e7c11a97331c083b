"""Bandwidth probes of the diagnostics library (need a GPU)."""
import pytest

from biotfd import lib, probes

pytestmark = pytest.mark.gpu


def test_copy_probe_reports_bandwidth():
    assert probes.probe_copy(1 << 22, 3) > 0


def test_march_probe_reports_bandwidth_and_exact_bytes():
    gbs, moved = probes.probe_march(1100, 300, 3)
    assert gbs > 0
    # 2 x 512 columns, 2 blocks of 128 rows: 9 arrays read over 132 rows, 8 written over 128
    assert moved == 1024 * 2 * 132 * 9 * 4 + 1024 * 2 * 128 * 8 * 4


def test_march_probe_rejects_small_grids():
    with pytest.raises(lib.BiotError):
        probes.probe_march(100, 300, 1)


def test_bands_probe_reports_bandwidth_and_rejects_bad_bands():
    assert probes.probe_bands(1100, 300, 1, 2) > 0
    assert probes.probe_bands(1100, 300, 7, 2) > 0   # 42 rows per band, 294 rows
    with pytest.raises(lib.BiotError):
        probes.probe_bands(1100, 300, 0, 1)
    with pytest.raises(lib.BiotError):
        probes.probe_bands(100, 300, 1, 1)
