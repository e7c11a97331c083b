"""Output writers with the reference's byte formats (serial build).

* :class:`SeismogramWriter` - STORE_SEISMOGRAMS (store_seismograms.F90:47-154)
  and STORE_SEISMOGRAMS_PRESSURE (store_seismograms_pressure.F90:26-81):
  ``REC_UW_<j>.asc`` (t, U, -W, dU/dt, -dW/dt), ``REC_Q_<j>.asc`` (the same for
  QX, QZ; with KEY_TLQ) and ``REC_STRESS_<j>.asc`` (t, PRES, -TXX; with
  KEY_TLP), one line per step in Fortran ``(n(1X,E13.6))`` format, appended.
  The reference opens and closes every file at every step; here a chunk of
  steps is written with one open per file (``bq_write_seismograms`` in the
  library's host code), which appends the same bytes.
* :func:`write_snapshots` - STORE_SNAPSHOTS (store_snapshots.F90): ``SNAP.V<step>``
  (UM, WM) and, with KEY_TLQ, ``SNAP.Q<step>`` (QX, QZ), one unformatted record each.

The samples are the six values per receiver and step that the library returns
(``bq_get_traces``): U, -W, QX, -QZ, PRES, -TXX at the receiver nodes. The time
derivatives are formed in binary32 as the reference forms them: the difference
to the previous sample over DT, the previous sample starting at 0. For
receivers outside the grid the reference writes 0 samples and 0 U/W
derivatives but leaves the QX/QZ derivative columns uninitialised
(store_seismograms.F90:37-45); they are written as 0 here.
"""
from __future__ import annotations

import os

import numpy as np

from . import lib
from .formats import write_snapshot


def fortran_e(x, w: int = 13, d: int = 6) -> str:
    """Fortran ``E13.6`` edit of a binary32 value as the reference's runtime
    writes it: 0.ddddddE+ee right-aligned; NaN, Inf, -Inf spelled out."""
    if (w, d) != (13, 6):
        raise ValueError("only the reference's E13.6 edit is provided")
    out = lib.C.create_string_buffer(14)
    if lib.load().bq_format_e13_6(np.float32(x), out):
        raise lib.BiotError("bq_format_e13_6 failed")
    return out.value.decode()


def fortran_line(values) -> str:
    return "".join(" " + fortran_e(v) for v in values) + "\n"


class SeismogramWriter:
    """Appends receiver samples to the reference's receiver files in
    ``directory``, a chunk of steps per :meth:`write_chunk` call."""

    def __init__(self, directory: str, dt: float, mr: int, key_tlq: bool = False,
                 key_tlp: bool = False, inside: np.ndarray | None = None):
        self.dir = directory
        self.dt = np.float32(dt)
        self.mr = mr
        self.key_tlq = key_tlq
        self.key_tlp = key_tlp
        # (mr, 2): whether the U node and the W node lie in the grid
        self.inside = None if inside is None else np.ascontiguousarray(inside, np.int32)
        self.state = np.zeros((mr, 4), np.float32)   # SSU, SSW, SSQX, SSQZ
        self.next_step = 1

    def write_chunk(self, first_step: int, samples: np.ndarray):
        """samples: (nsteps, mr, 6) binary32 of steps first_step .. first_step+nsteps-1."""
        s = np.ascontiguousarray(samples, np.float32)
        if s.ndim != 3 or s.shape[1:] != (self.mr, 6):
            raise ValueError(f"samples of shape {s.shape}, need (nsteps, {self.mr}, 6)")
        if first_step != self.next_step:
            raise ValueError(f"steps must follow one another: next is {self.next_step}")
        ins = self.inside.ctypes.data_as(lib.C.c_void_p) if self.inside is not None else None
        rc = lib.load().bq_write_seismograms(
            self.dir.encode(), s.ctypes.data_as(lib.C.c_void_p), first_step, s.shape[0], self.mr,
            self.dt, ins, int(self.key_tlq), int(self.key_tlp),
            self.state.ctypes.data_as(lib.C.c_void_p))
        if rc:
            raise lib.BiotError(f"bq_write_seismograms failed ({rc})")
        self.next_step = first_step + s.shape[0]

    def write(self, step: int, samples: np.ndarray):
        """One step: samples (mr, 6)."""
        self.write_chunk(step, np.asarray(samples, np.float32)[None])


def receivers_inside(receivers, pmll: int, pmlxh: int, pmlzh: int) -> np.ndarray:
    """(mr, 2) bool: U node (IRECX, LRECX) and W node (IRECZ, LRECZ) inside the
    grid (store_seismograms.F90:51-52, 79-80)."""
    r = np.asarray(receivers, np.int64).reshape(-1, 4)
    ok_u = (r[:, 0] >= pmll) & (r[:, 0] <= pmlxh) & (r[:, 1] >= 0) & (r[:, 1] <= pmlzh)
    ok_w = (r[:, 2] >= pmll) & (r[:, 2] <= pmlxh) & (r[:, 3] >= 0) & (r[:, 3] <= pmlzh)
    return np.stack([ok_u, ok_w], axis=1)


def write_snapshots(directory: str, step: int, u, w, qx=None, qz=None):
    """SNAP.V<step> (and SNAP.Q<step> when qx, qz are given)."""
    write_snapshot(os.path.join(directory, f"SNAP.V{step:06d}"), u, w)
    if qx is not None:
        write_snapshot(os.path.join(directory, f"SNAP.Q{step:06d}"), qx, qz)
