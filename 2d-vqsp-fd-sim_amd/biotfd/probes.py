"""ctypes binding of the bandwidth probes (``include/biotfd_probes.h``,
``libbiotfd_probes.so``). Diagnostics only: they are not part of the time
loop and have no counterpart in the reference."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .lib import BiotError

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libbiotfd_probes.so")
EXPORTS = ("bq_probe_copy", "bq_probe_march", "bq_probe_bands")

_lib = None


def load(path: str = LIBPATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise BiotError(f"{path} is not built; run python 2d-vqsp-fd-sim_amd/build.py")
    lib = C.CDLL(path)
    p = C.c_void_p
    lib.bq_probe_copy.argtypes = [C.c_longlong, C.c_int, p]
    lib.bq_probe_copy.restype = C.c_int
    lib.bq_probe_march.argtypes = [C.c_int, C.c_int, C.c_int, p, p]
    lib.bq_probe_march.restype = C.c_int
    lib.bq_probe_bands.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, p]
    lib.bq_probe_bands.restype = C.c_int
    _lib = lib
    return lib


def probe_copy(nfloats: int, iters: int = 20) -> float:
    """Read+write GB/s of a plain device-to-device copy of nfloats floats."""
    out = np.zeros(1, dtype=np.float64)
    rc = load().bq_probe_copy(int(nfloats), int(iters), out.ctypes.data_as(C.c_void_p))
    if rc:
        raise BiotError(f"bq_probe_copy failed ({rc})")
    return float(out[0])


def probe_march(nx: int, nz: int, iters: int = 20):
    """(GB/s, bytes per launch) of the fused tiles' row-streaming access pattern
    (nine arrays read, eight written, no arithmetic) on an nx x nz grid."""
    out = np.zeros(2, dtype=np.float64)
    rc = load().bq_probe_march(int(nx), int(nz), int(iters), out.ctypes.data_as(C.c_void_p),
                               out[1:].ctypes.data_as(C.c_void_p))
    if rc:
        raise BiotError(f"bq_probe_march failed ({rc})")
    return float(out[0]), float(out[1])


def probe_bands(nx: int, nz: int, bands: int, iters: int = 20) -> float:
    """GB/s of the fused tiles' per-cell traffic (nine arrays read, eight
    written) without the march, with the blocks in flight over `bands` row
    bands (1: row-major)."""
    out = np.zeros(1, dtype=np.float64)
    rc = load().bq_probe_bands(int(nx), int(nz), int(bands), int(iters), out.ctypes.data_as(C.c_void_p))
    if rc:
        raise BiotError(f"bq_probe_bands failed ({rc})")
    return float(out[0])
