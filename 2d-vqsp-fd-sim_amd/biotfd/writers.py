"""Output writers with the reference's byte formats (serial build).

* :class:`SeismogramWriter` - STORE_SEISMOGRAMS (store_seismograms.F90:47-154)
  and STORE_SEISMOGRAMS_PRESSURE (store_seismograms_pressure.F90:26-81):
  ``REC_UW_<j>.asc`` (t, U, -W, dU/dt, -dW/dt), ``REC_Q_<j>.asc`` (the same for
  QX, QZ; with KEY_TLQ) and ``REC_STRESS_<j>.asc`` (t, PRES, -TXX; with
  KEY_TLP), one line per step in Fortran ``(n(1X,E13.6))`` format, appended.
* :func:`write_snapshots` - STORE_SNAPSHOTS (store_snapshots.F90): ``SNAP.V<step>``
  (UM, WM) and, with KEY_TLQ, ``SNAP.Q<step>`` (QX, QZ), one unformatted record each.

The samples are the six values per receiver and step that the library returns
(``bq_get_traces``): U, -W, QX, -QZ, PRES, -TXX at the receiver nodes. The time
derivatives are formed here in binary32 as the reference forms them: the
difference to the previous sample over DT, the previous sample starting at 0.
For receivers outside the grid the reference writes 0 samples and 0 U/W
derivatives but leaves the QX/QZ derivative columns uninitialised
(store_seismograms.F90:37-45); they are written as 0 here.
"""
from __future__ import annotations

import math
import os

import numpy as np

from .formats import write_snapshot


def fortran_e(x, w: int = 13, d: int = 6) -> str:
    """Fortran ``Ew.d`` edit of a binary32 value: 0.ddddddE+ee, right-aligned."""
    v = float(np.float32(x))
    neg = math.copysign(1.0, v) < 0
    if v == 0.0:
        digits, exp = "0" * d, 0
    else:
        s = f"{abs(v):.{d - 1}e}"          # d significant digits, correctly rounded
        mant, e = s.split("e")
        digits = mant.replace(".", "")
        exp = int(e) + 1
    body = f"{'-' if neg else ''}0.{digits}E{'+' if exp >= 0 else '-'}{abs(exp):02d}"
    return body.rjust(w)


def fortran_line(values) -> str:
    return "".join(" " + fortran_e(v) for v in values) + "\n"


class SeismogramWriter:
    """Appends one step of receiver samples per :meth:`write` call to the
    reference's receiver files in ``directory``."""

    def __init__(self, directory: str, dt: float, mr: int, key_tlq: bool = False,
                 key_tlp: bool = False):
        self.dir = directory
        self.dt = np.float32(dt)
        self.mr = mr
        self.key_tlq = key_tlq
        self.key_tlp = key_tlp
        self.prev = np.zeros((mr, 4), np.float32)   # SSU, SSW, SSQX, SSQZ

    def write(self, step: int, samples: np.ndarray, inside: np.ndarray | None = None):
        """samples: (mr, 6) binary32 of one step; inside: (mr, 2) bool, whether
        the U node and the W node lie in the grid (default: all)."""
        s = np.asarray(samples, np.float32)
        if inside is None:
            inside = np.ones((self.mr, 2), bool)
        t = np.float32(self.dt * np.float32(step))
        with np.errstate(all="ignore"):
            for j in range(self.mr):
                d = np.zeros(4, np.float32)
                for c, (k, node) in enumerate(((0, 0), (1, 1), (2, 0), (3, 1))):
                    if inside[j, node]:
                        d[c] = np.float32(np.float32(s[j, k] - self.prev[j, c]) / self.dt)
                        self.prev[j, c] = s[j, k]
                name = str(j + 1)
                with open(os.path.join(self.dir, f"REC_UW_{name}.asc"), "a") as f:
                    f.write(fortran_line((t, self.prev[j, 0], self.prev[j, 1], d[0], d[1])))
                if self.key_tlq:
                    with open(os.path.join(self.dir, f"REC_Q_{name}.asc"), "a") as f:
                        f.write(fortran_line((t, self.prev[j, 2], self.prev[j, 3], d[2], d[3])))
                if self.key_tlp:
                    with open(os.path.join(self.dir, f"REC_STRESS_{name}.asc"), "a") as f:
                        f.write(fortran_line((t, s[j, 4], s[j, 5])))


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
