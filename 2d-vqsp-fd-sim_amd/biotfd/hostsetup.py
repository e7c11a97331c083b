"""Host-side set-up that the reference performs before TIME_LOOP.

In the drop-in these steps stay in the reference's Fortran driver; they are
restated here so that Python callers (tests, the bench, the batching front-end)
can prepare a run exactly as the Fortran driver does, in the same arithmetic.

* :func:`jm_grid` - READ_MODEL (``read_model.F90:57-87``): material indices on
  the computational grid, replicated laterally into the PML, into the two
  image rows and into the bottom PML rows.
* :func:`scaled_table` - READ_MODEL_PAR scaling (``read_model_par.F90:63-74``)
  in the working precision.
* :func:`read_stf_reals` - the ``STF.DAT`` samples as a ``REAL(PP)`` read sees them.
"""
from __future__ import annotations

import numpy as np

from .formats import Job, MO_INDEX


def grid_shape(job: Job):
    """(NZc, NXc) of the computational grid (aux1.F90:36-46, read_open.F90:149-151)."""
    nx = job.mx + 2 * job.tpml
    nz = job.mzp + job.tpml + 1
    return nz, nx


def jm_grid(job: Job) -> np.ndarray:
    """READ_MODEL for the serial build: int32 ``[NZc][NXc]``, x fastest."""
    T, mx, mzp = job.tpml, job.mx, job.mzp
    nz, nx = grid_shape(job)
    jm = np.zeros((nz, nx), dtype=np.int32)
    phys = np.asarray(job.jm, dtype=np.int32)
    first = 2 if job.stress_imaging else 0
    if phys.shape != (mzp - first + 1, mx):
        raise ValueError(f"jm must have shape {(mzp - first + 1, mx)}, got {phys.shape}")
    # rows first..MZ' hold the file records in columns 1..MX (x index i -> column i+T-1)
    jm[first:mzp + 1, T:T + mx] = phys
    for L in range(first, mzp + 1):
        jm[L, :T] = jm[L, T]                 # JMT(1-J, L) = JMT(1, L), J = 0..T
        jm[L, T + mx:] = jm[L, T + mx - 1]   # JMT(MX+J, L) = JMT(MX, L)
    if job.stress_imaging:
        jm[1, :] = jm[2, :]                  # JMT(:,1) = JMT(:,2)
        jm[0, :] = jm[3, :]                  # JMT(:,0) = JMT(:,3)
    jm[mzp + 1:, :] = jm[mzp, :]             # JMT(:,L) = JMT(:,MZ), L = MZ+1..PMLZH
    return jm


def scaled_table(job: Job, dtype=np.float32) -> np.ndarray:
    """Material table after READ_MODEL_PAR: ``[JMNUM][24]`` in ``dtype``.

    The raw table is first rounded to ``dtype`` (as the ``.MO`` record holds
    ``REAL(PP)`` values), then REL* entries are scaled by DT/H or DT with the
    reference's expression order ``DT*X/H``.
    """
    t = np.asarray(job.table, dtype=np.float64).astype(dtype).copy()
    dt = dtype(job.dt)
    h = dtype(job.h)
    over_h = ("REL1_U", "REL3_U", "REL1_W", "REL3_W", "REL1_QX", "REL2_QX", "REL1_QZ", "REL2_QZ")
    plain = ("REL2_U", "REL2_W", "REL3_QX", "REL3_QZ")
    for k in over_h:
        c = MO_INDEX[k]
        t[:, c] = (dt * t[:, c]) / h
    for k in plain:
        c = MO_INDEX[k]
        t[:, c] = dt * t[:, c]
    return t


def read_stf_reals(job: Job, dtype=np.float32):
    """``STF.DAT`` samples in the working precision."""
    return (np.asarray(job.stf_t, dtype=np.float32).astype(dtype),
            np.asarray(job.stf_v, dtype=np.float32).astype(dtype))


def nr_params(job: Job, dtype=np.float32) -> np.ndarray:
    """OMG, WB, THPPLE, THSSLE as MODULE CONTROL_DATA holds them after
    READ_OPEN: the namelist OMG divided by the default-real literal 2 pi
    (read_open.F90:188), in the working precision."""
    two_pi = dtype(np.float32(6.283185307179586476925286766559))
    omg = dtype(job.omg) / two_pi
    return np.array([omg, job.wb, job.thpple, job.thssle], dtype=dtype)
