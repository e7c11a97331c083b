"""ctypes binding of the C ABI in ``include/biotfd.h`` (``libbiotfd.so``).

The library is the HIP implementation; there is no CPU fallback. Loading fails
loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import hostsetup
from .formats import Job

HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build; BIOTFD_LIB names another build of the same library
LIBPATH = os.environ.get("BIOTFD_LIB") or os.path.join(HERE, "libbiotfd.so")

FIELDS = ("U", "W", "QX", "QZ", "TXX", "TZZ", "TXZ", "P")
EXPORTS = (
    "bq_create", "bq_destroy", "bq_upload_model", "bq_build_pml", "bq_upload_pml", "bq_get_pml",
    "bq_upload_bottom_nr", "bq_set_source", "bq_set_receivers", "bq_run", "bq_sync",
    "bq_get_traces", "bq_get_field", "bq_set_field", "bq_source_from_stf", "bq_last_error",
    "bq_grid_shape", "bq_enable_timing", "bq_kernel_times", "bq_kernel_cells",
    "bq_set_sweep", "bq_get_sweep", "bq_kernel_stats", "bq_set_tiles", "bq_get_tiles",
    "bq_autotune", "bq_get_tile_rows",
    "bq_build_bottom_nr", "bq_nonref", "bq_write_seismograms", "bq_format_e13_6",
    "bq_batch_create", "bq_batch_destroy", "bq_batch_run", "bq_batch_sync", "bq_batch_last_run_ms",
    "bq_batch_autotune", "bq_batch_set_schedule", "bq_batch_get_schedule", "bq_batch_last_error",
    "bq_trace_misfit", "bq_trace_misfit_noalloc",
)

SCHED_GRID, SCHED_PERSISTENT = 0, 1
# bq_set_tiles modes, and the candidates of bq_autotune in order
TILES = {"auto": 0, "fused": 1, "cells": 2}
TUNE_TILES = ("fused128", "fused64", "fused48", "fused32", "fused24", "fused16", "cells")
# candidate c of bq_batch_autotune: tile height TUNE_ROWS[c // 2], schedule c % 2
TUNE_ROWS = (64, 128, 256)


class Desc(C.Structure):
    _fields_ = [("mx", C.c_int32), ("mz", C.c_int32), ("tpml", C.c_int32),
                ("stress_imaging", C.c_int32), ("lin_x", C.c_int32), ("lin_z", C.c_int32),
                ("ktbo", C.c_int32), ("mt2", C.c_int32), ("jmnum", C.c_int32),
                ("variant", C.c_int32), ("h", C.c_double), ("dt", C.c_double)]


class BiotError(RuntimeError):
    pass


_lib = None


def load(path: str = LIBPATH):
    """Load ``libbiotfd.so``; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise BiotError(f"{path} is not built; run __graft_entry__.build() or "
                        "python 2d-vqsp-fd-sim_amd/build.py")
    lib = C.CDLL(path)
    p = C.c_void_p
    sig = {
        "bq_create": ([C.POINTER(Desc), C.POINTER(p)], C.c_int),
        "bq_destroy": ([p], None),
        "bq_upload_model": ([p, p, p], C.c_int),
        "bq_build_pml": ([p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float], C.c_int),
        "bq_upload_pml": ([p, p], C.c_int),
        "bq_get_pml": ([p, C.c_int, p], C.c_int),
        "bq_upload_bottom_nr": ([p, p, p], C.c_int),
        "bq_set_source": ([p, p], C.c_int),
        "bq_set_receivers": ([p, C.c_int, p, p, p, p], C.c_int),
        "bq_run": ([p, C.c_int, C.c_int], C.c_int),
        "bq_sync": ([p], C.c_int),
        "bq_get_traces": ([p, p, C.c_int, C.c_int], C.c_int),
        "bq_get_field": ([p, C.c_int, p], C.c_int),
        "bq_set_field": ([p, C.c_int, p], C.c_int),
        "bq_source_from_stf": ([p, p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.c_float, p],
                               C.c_int),
        "bq_last_error": ([p], C.c_char_p),
        "bq_grid_shape": ([p, p, p], C.c_int),
        "bq_enable_timing": ([p, C.c_int], C.c_int),
        "bq_kernel_times": ([p, p], C.c_int),
        "bq_kernel_cells": ([p, p], C.c_int),
        "bq_set_sweep": ([p, C.c_int], C.c_int),
        "bq_get_sweep": ([p], C.c_int),
        "bq_set_tiles": ([p, C.c_int], C.c_int),
        "bq_get_tiles": ([p], C.c_int),
        "bq_autotune": ([p, C.c_int, p], C.c_int),
        "bq_get_tile_rows": ([p], C.c_int),
        "bq_kernel_stats": ([p, p, p, p], C.c_int),
        "bq_build_bottom_nr": ([p, p, p, p], C.c_int),
        "bq_nonref": ([C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, p, p],
                      C.c_int),
        "bq_write_seismograms": ([C.c_char_p, p, C.c_int, C.c_int, C.c_int, C.c_float, p, C.c_int,
                                  C.c_int, p], C.c_int),
        "bq_format_e13_6": ([C.c_float, C.c_char_p], C.c_int),
        "bq_batch_create": ([p, C.c_int, C.POINTER(p)], C.c_int),
        "bq_batch_destroy": ([p], None),
        "bq_batch_run": ([p, C.c_int], C.c_int),
        "bq_batch_sync": ([p], C.c_int),
        "bq_batch_last_run_ms": ([p, p], C.c_int),
        "bq_batch_autotune": ([p, C.c_int, p], C.c_int),
        "bq_batch_set_schedule": ([p, C.c_int, C.c_int, C.c_int], C.c_int),
        "bq_batch_get_schedule": ([p, p, p, p], C.c_int),
        "bq_batch_last_error": ([p], C.c_char_p),
        "bq_trace_misfit": ([p, p, C.c_int, C.c_int, p, p], C.c_int),
        "bq_trace_misfit_noalloc": ([p, p, C.c_int, C.c_int, p, p], C.c_int),
    }
    for name, (args, res) in sig.items():
        try:
            f = getattr(lib, name)
        except AttributeError:
            raise BiotError(f"{path} does not export {name}") from None
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def nonref(key: int, h, dt, vel, vels, velp, prm) -> np.ndarray:
    """NONREF (nonref.F90:56-206) for one boundary: AX01 AX02 AX10 AX11 AX12
    AX20 AX21 AX22, computed by the library's host code."""
    prm = np.ascontiguousarray(prm, np.float32)
    out = np.zeros(8, np.float32)
    rc = load().bq_nonref(int(key), float(h), float(dt), float(vel), float(vels), float(velp),
                          _ptr(prm), _ptr(out))
    if rc:
        raise BiotError(f"bq_nonref failed ({rc})")
    return out


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def source_from_stf(job: Job) -> np.ndarray:
    """OPEN_SR_FILE through the library's host helper: SOURTF_A as [MT2][2]."""
    lib = load()
    t, v = hostsetup.read_stf_reals(job, np.float32)
    out = np.zeros((job.mt2, 2), dtype=np.float32)
    rc = lib.bq_source_from_stf(_ptr(t), _ptr(v), len(t), job.mt2, np.float32(job.dt),
                                np.float32(job.h), job.key_sour, np.float32(job.qr), _ptr(out))
    if rc:
        raise BiotError(f"bq_source_from_stf failed ({rc})")
    return out


class Simulation:
    """One run on the GPU, driven like the reference's TIME_LOOP.

    Set-up follows the reference driver's order (main.F90:60-84): model
    (READ_MODEL, READ_MODEL_PAR), source (OPEN_SR_FILE), bottom coefficients
    (PRE_NR_COEF1), PML (PPML); then :meth:`run` advances time steps.
    """

    def __init__(self, job: Job, abw=None, abh=None, pml_arrays=None, sourtf=None):
        self.lib = load()
        self.job = job
        self.nz, self.nx = hostsetup.grid_shape(job)
        d = Desc(job.mx, job.mzp, job.tpml, int(job.stress_imaging), job.lin_x, job.lin_z,
                 job.ktbo, job.mt2, job.table.shape[0], int(job.variant), job.h, float(np.float32(job.dt)))
        h = C.c_void_p()
        rc = self.lib.bq_create(C.byref(d), C.byref(h))
        if rc:
            raise BiotError(f"bq_create failed ({rc})")
        self.h = h
        self.jm = hostsetup.jm_grid(job)
        self.tab = np.ascontiguousarray(hostsetup.scaled_table(job, np.float32))
        self._chk(self.lib.bq_upload_model(h, _ptr(self.jm), _ptr(self.tab)))
        if pml_arrays is None:
            self._chk(self.lib.bq_build_pml(h, job.r, job.tau, job.pow, job.sf, job.wc))
        else:
            arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in pml_arrays]
            ptrs = (C.c_void_p * 24)(*[a.ctypes.data for a in arrs])
            self._chk(self.lib.bq_upload_pml(h, ptrs))
        if abw is None and abh is None and job.ktbo != 0:
            # PRE_NR_COEF1's bottom coefficients, computed by the library
            abw = np.zeros(8, np.float32)
            abh = np.zeros(8, np.float32)
            prm = hostsetup.nr_params(job, np.float32)
            self._chk(self.lib.bq_build_bottom_nr(h, _ptr(prm), _ptr(abw), _ptr(abh)))
        else:
            abw = np.zeros(8, np.float32) if abw is None else np.asarray(abw, np.float32)
            abh = np.zeros(8, np.float32) if abh is None else np.asarray(abh, np.float32)
            self._chk(self.lib.bq_upload_bottom_nr(h, _ptr(abw), _ptr(abh)))
        self.abw, self.abh = abw, abh
        self.sourtf = source_from_stf(job) if sourtf is None else np.ascontiguousarray(sourtf, np.float32)
        self._chk(self.lib.bq_set_source(h, _ptr(self.sourtf)))
        rec = np.asarray(job.receivers, dtype=np.int32).reshape(-1, 4)
        self.mr = rec.shape[0]
        if self.mr:
            cols = [np.ascontiguousarray(rec[:, k]) for k in range(4)]
            self._chk(self.lib.bq_set_receivers(h, self.mr, *[_ptr(c) for c in cols]))
        self.step = 0

    def _chk(self, rc):
        if rc:
            msg = self.lib.bq_last_error(self.h)
            raise BiotError(f"biotfd error {rc}: {msg.decode() if msg else ''}")

    def run(self, nsteps: int, sync: bool = True):
        self._chk(self.lib.bq_run(self.h, self.step + 1, self.step + nsteps))
        self.step += nsteps
        if sync:
            self.sync()

    def sync(self):
        self._chk(self.lib.bq_sync(self.h))

    def enable_timing(self, every: int = 1):
        """Events around the launches of every ``every``-th step of each run
        (0: off)."""
        self._chk(self.lib.bq_enable_timing(self.h, int(every)))

    def kernel_times(self) -> dict:
        """ms per step of each launch: k_step (fused interior + boundary stress +
        previous step's samples; k_rest_stress without fused tiles), k_velocity
        (boundary velocity, with the edge rows where its edge blocks take them),
        k_edges (0 then)."""
        out = np.zeros(3, dtype=np.float64)
        self._chk(self.lib.bq_kernel_times(self.h, _ptr(out)))
        return dict(zip(("k_step", "k_velocity", "k_edges"), map(float, out)))

    def kernel_cells(self) -> dict:
        """Cells updated per launch of k_fused, k_stress and k_velocity."""
        out = np.zeros(3, dtype=np.int64)
        self._chk(self.lib.bq_kernel_cells(self.h, _ptr(out)))
        return dict(zip(("k_fused", "k_stress", "k_velocity"), map(int, out)))

    def set_sweep(self, steps: int):
        """Steps per sweep bq_run may use: 1 (the default) or 2 (where the
        grid allows)."""
        self._chk(self.lib.bq_set_sweep(self.h, int(steps)))

    def sweep(self) -> int:
        """Steps per sweep the next run uses (bq_get_sweep)."""
        return int(self.lib.bq_get_sweep(self.h))

    def set_tiles(self, mode: str):
        """Fused tiles or not (bq_set_tiles): "auto" (the default), "fused" or
        "cells"."""
        self._chk(self.lib.bq_set_tiles(self.h, TILES[mode]))

    def tiles(self) -> str:
        """Partition the next run uses: "fused" or "cells" (bq_get_tiles)."""
        return {1: "fused", 2: "cells"}[int(self.lib.bq_get_tiles(self.h))]

    def tile_rows(self) -> int:
        """Rows per fused block of the partition with fused tiles (bq_get_tile_rows)."""
        return int(self.lib.bq_get_tile_rows(self.h))

    def autotune(self, trial_steps: int = 4) -> dict:
        """Runs len(TUNE_TILES) * (trial_steps + 1) real steps, one set per
        partition, and keeps the fastest (bq_autotune). Returns ms per step by
        candidate: fused blocks of 128 .. 16 rows, then "cells"."""
        ms = np.zeros(len(TUNE_TILES), np.float64)
        self._chk(self.lib.bq_autotune(self.h, int(trial_steps), _ptr(ms)))
        self.step += len(TUNE_TILES) * (int(trial_steps) + 1)
        return dict(zip(TUNE_TILES, map(float, ms)))

    def kernel_stats(self) -> dict:
        """Timing of the last timed run by launch kind (bq_kernel_stats)."""
        ms = np.zeros(4, np.float64)
        cells = np.zeros(4, np.int64)
        sw = np.zeros(1, np.int32)
        self._chk(self.lib.bq_kernel_stats(self.h, _ptr(ms), _ptr(cells), _ptr(sw)))
        if sw[0] == 2:
            names = ("k_sweep2", "k_rest_stress", "k_velocity", "k_edges")
            cnames = ("region", "stress", "velocity")
        else:
            names = ("k_step", "k_velocity", "k_edges")
            cnames = ("k_fused", "k_stress", "k_velocity")
        return {"steps_per_sweep": int(sw[0]), "ms": dict(zip(names, map(float, ms))),
                "cells": dict(zip(cnames, map(int, cells)))}

    def field(self, name: str) -> np.ndarray:
        out = np.zeros((self.nz, self.nx), dtype=np.float32)
        self._chk(self.lib.bq_get_field(self.h, FIELDS.index(name), _ptr(out)))
        return out

    def set_field(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float32)
        self._chk(self.lib.bq_set_field(self.h, FIELDS.index(name), _ptr(a)))

    def traces(self, first: int = 1, n: int | None = None) -> np.ndarray:
        n = self.step - first + 1 if n is None else n
        out = np.zeros((n, self.mr, 6), dtype=np.float32)
        self._chk(self.lib.bq_get_traces(self.h, _ptr(out), first, n))
        return out

    def pml(self, k: int) -> np.ndarray:
        n = self.lib.bq_get_pml(self.h, k, None)
        out = np.zeros(n, dtype=np.float32)
        self._chk(min(0, self.lib.bq_get_pml(self.h, k, _ptr(out))))
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.bq_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def trace_misfit(a: "Simulation", b: "Simulation", first: int = 1, n: int | None = None,
                 per_receiver: bool = False, entry: str = "bq_trace_misfit"):
    """Sums of the trace misfit of run a against run b on the device
    (bq_trace_misfit; entry="bq_trace_misfit_noalloc" for the variant that
    allocates nothing): (sum (a-b)^2, sum b^2) over steps first .. first+n-1,
    and per receiver [mr, 2] when per_receiver is set."""
    n = min(a.step, b.step) - first + 1 if n is None else n
    tot = np.zeros(2, np.float64)
    per = np.zeros((a.mr, 2), np.float64) if per_receiver else None
    rc = getattr(load(), entry)(a.h, b.h, int(first), int(n),
                                _ptr(per) if per is not None else None, _ptr(tot))
    a._chk(rc)
    return (tot, per) if per_receiver else tot


class Batch:
    """Many :class:`Simulation` shots advanced together (``bq_batch_*``): three
    launches per time step for all of them, replayed from a HIP graph. Each
    shot's results are bit-identical to running it alone; read them from the
    shots as usual."""

    def __init__(self, sims):
        self.lib = load()
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s.h.value for s in self.sims])
        b = C.c_void_p()
        rc = self.lib.bq_batch_create(arr, len(self.sims), C.byref(b))
        if rc:
            msg = self.lib.bq_batch_last_error(None)
            raise BiotError(f"bq_batch_create failed ({rc}): {msg.decode() if msg else ''}")
        self.b = b

    def _chk(self, rc):
        if rc:
            msg = self.lib.bq_batch_last_error(self.b)
            raise BiotError(f"biotfd batch error {rc}: {msg.decode() if msg else ''}")

    def run(self, nsteps: int, sync: bool = True):
        self._chk(self.lib.bq_batch_run(self.b, int(nsteps)))
        for s in self.sims:
            s.step += nsteps
        if sync:
            self.sync()

    def sync(self):
        self._chk(self.lib.bq_batch_sync(self.b))

    def last_run_ms(self) -> float:
        out = np.zeros(1, np.float64)
        self._chk(self.lib.bq_batch_last_run_ms(self.b, _ptr(out)))
        return float(out[0])

    def tile_rows(self) -> int:
        """Rows per fused block of the partition with fused tiles (bq_get_tile_rows)."""
        return int(self.lib.bq_get_tile_rows(self.h))

    def autotune(self, trial_steps: int = 4) -> dict:
        """Times the six candidates on the batch's own steps (6 * trial_steps
        steps, which advance every shot) and keeps the fastest; returns
        {(tile_rows, schedule): ms per step}."""
        out = np.zeros(6, np.float64)
        self._chk(self.lib.bq_batch_autotune(self.b, int(trial_steps), _ptr(out)))
        for s in self.sims:
            s.step += 6 * trial_steps
        return {(TUNE_ROWS[c // 2], c % 2): float(out[c]) for c in range(6)}

    def set_schedule(self, tile_rows: int = 128, schedule: int = SCHED_GRID, graph: bool = True):
        self._chk(self.lib.bq_batch_set_schedule(self.b, int(tile_rows), int(schedule), int(graph)))

    def schedule(self):
        out = np.zeros(3, np.int32)
        self._chk(self.lib.bq_batch_get_schedule(self.b, _ptr(out), _ptr(out[1:]), _ptr(out[2:])))
        return int(out[0]), int(out[1]), bool(out[2])

    def close(self):
        if getattr(self, "b", None):
            self.lib.bq_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
