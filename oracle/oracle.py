"""ctypes binding of the CPU oracle (test infrastructure only).

Loads ``liboracle_f32.so`` / ``liboracle_f64.so`` built by ``oracle/Makefile``.
Only tests, the smoke check and the CPU-baseline leg of the bench import this.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "2d-vqsp-fd-sim_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

from biotfd import hostsetup  # noqa: E402
from biotfd.formats import Job  # noqa: E402

FIELDS = ("U", "W", "QX", "QZ", "TXX", "TZZ", "TXZ", "P", "QXP", "QZP")


class Desc(C.Structure):
    _fields_ = [("mx", C.c_int32), ("mz", C.c_int32), ("tpml", C.c_int32),
                ("stress_imaging", C.c_int32), ("lin_x", C.c_int32), ("lin_z", C.c_int32),
                ("ktbo", C.c_int32), ("mt2", C.c_int32), ("jmnum", C.c_int32),
                ("variant", C.c_int32), ("h", C.c_double), ("dt", C.c_double)]


_LIBS = {}


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(precision: str = "f32"):
    if precision in _LIBS:
        return _LIBS[precision]
    path = os.path.join(HERE, f"liboracle_{precision}.so")
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    p = C.c_void_p
    lib.oracle_create.argtypes = [C.POINTER(Desc), C.POINTER(p)]
    for name, args in (
        ("oracle_destroy", [p]),
        ("oracle_upload_model", [p, p, p]),
        ("oracle_build_pml", [p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]),
        ("oracle_set_bottom_nr", [p, p, p]),
        ("oracle_set_source", [p, p]),
        ("oracle_set_receivers", [p, C.c_int, p, p, p, p]),
        ("oracle_run", [p, C.c_int, C.c_int]),
        ("oracle_get_field", [p, C.c_int, p]),
        ("oracle_set_field", [p, C.c_int, p]),
        ("oracle_get_traces", [p, p, C.c_int, C.c_int]),
        ("oracle_get_pml", [p, C.c_int, p]),
        ("oracle_bottom_nr", [p, C.c_int, p, p, p]),
        ("oracle_source_from_stf", [p, p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                    C.c_double, p]),
    ):
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = None if name == "oracle_destroy" else C.c_int
    real = C.c_float if precision == "f32" else C.c_double
    lib.oracle_nonref.argtypes = [C.c_int, real, real, real, real, real, p, p]
    lib.oracle_nonref.restype = C.c_int
    lib.oracle_last_error.argtypes = [p]
    lib.oracle_last_error.restype = C.c_char_p
    lib.oracle_real_size.restype = C.c_int
    _LIBS[precision] = lib
    return lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def source_from_stf(job: Job, precision="f32"):
    lib = load(precision)
    dt = np.float32 if precision == "f32" else np.float64
    t, v = hostsetup.read_stf_reals(job, dt)
    out = np.zeros((job.mt2, 2), dtype=dt)
    rc = lib.oracle_source_from_stf(_ptr(t), _ptr(v), len(t), job.mt2, float(np.float32(job.dt)) if precision == "f32" else job.dt,
                                    job.h, job.key_sour, job.qr, _ptr(out))
    if rc:
        raise RuntimeError(f"oracle_source_from_stf failed ({rc})")
    return out


class OracleSim:
    """One run of the CPU oracle, driven like the reference's TIME_LOOP."""

    def __init__(self, job: Job, precision: str = "f32", abw=None, abh=None):
        self.lib = load(precision)
        self.dtype = np.float32 if precision == "f32" else np.float64
        self.job = job
        self.nz, self.nx = hostsetup.grid_shape(job)
        d = Desc(job.mx, job.mzp, job.tpml, int(job.stress_imaging), job.lin_x, job.lin_z,
                 job.ktbo, job.mt2, job.table.shape[0], int(job.variant), job.h,
                 float(np.float32(job.dt)) if precision == "f32" else job.dt)
        h = C.c_void_p()
        rc = self.lib.oracle_create(C.byref(d), C.byref(h))
        if rc:
            raise RuntimeError(f"oracle_create failed ({rc})")
        self.h = h
        self.jm = hostsetup.jm_grid(job)
        self.tab = np.ascontiguousarray(hostsetup.scaled_table(job, self.dtype))
        self._chk(self.lib.oracle_upload_model(h, _ptr(self.jm), _ptr(self.tab)))
        self._chk(self.lib.oracle_build_pml(h, job.r, job.tau, job.pow, job.sf, job.wc))
        if abw is None and abh is None and job.ktbo != 0:
            abw = np.zeros(8, self.dtype)
            abh = np.zeros(8, self.dtype)
            prm = hostsetup.nr_params(job, self.dtype)
            self._chk(self.lib.oracle_bottom_nr(h, job.ktbo, _ptr(prm), _ptr(abw), _ptr(abh)))
        abw = np.zeros(8, self.dtype) if abw is None else np.asarray(abw, self.dtype)
        abh = np.zeros(8, self.dtype) if abh is None else np.asarray(abh, self.dtype)
        self.abw, self.abh = abw, abh
        self._chk(self.lib.oracle_set_bottom_nr(h, _ptr(abw), _ptr(abh)))
        self.sourtf = source_from_stf(job, precision)
        self._chk(self.lib.oracle_set_source(h, _ptr(self.sourtf)))
        rec = np.asarray(job.receivers, dtype=np.int32).reshape(-1, 4)
        self.mr = rec.shape[0]
        if self.mr:
            cols = [np.ascontiguousarray(rec[:, k]) for k in range(4)]
            self._chk(self.lib.oracle_set_receivers(h, self.mr, *[_ptr(c) for c in cols]))
        self.step = 0

    def _chk(self, rc):
        if rc:
            raise RuntimeError(f"oracle error {rc}: {self.lib.oracle_last_error(self.h)}")

    def run(self, nsteps: int):
        self._chk(self.lib.oracle_run(self.h, self.step + 1, self.step + nsteps))
        self.step += nsteps

    def field(self, name: str) -> np.ndarray:
        out = np.zeros((self.nz, self.nx), dtype=self.dtype)
        self._chk(self.lib.oracle_get_field(self.h, FIELDS.index(name), _ptr(out)))
        return out

    def set_field(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        self._chk(self.lib.oracle_set_field(self.h, FIELDS.index(name), _ptr(a)))

    def traces(self, first=1, n=None) -> np.ndarray:
        n = self.step - first + 1 if n is None else n
        out = np.zeros((n, self.mr, 6), dtype=self.dtype)
        self._chk(self.lib.oracle_get_traces(self.h, _ptr(out), first, n))
        return out

    def pml(self, k: int) -> np.ndarray:
        n = self.lib.oracle_get_pml(self.h, k, None)
        out = np.zeros(n, dtype=self.dtype)
        self.lib.oracle_get_pml(self.h, k, _ptr(out))
        return out

    def close(self):
        if self.h:
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def nonref(key, h, dt, vel, vels, velp, prm, precision="f32"):
    """NONREF (nonref.F90:56-206) for one boundary: the eight coefficients
    AX01, AX02, AX10, AX11, AX12, AX20, AX21, AX22."""
    lib = load(precision)
    dtype = np.float32 if precision == "f32" else np.float64
    prm = np.ascontiguousarray(prm, dtype)
    out = np.zeros(8, dtype)
    rc = lib.oracle_nonref(int(key), h, dt, vel, vels, velp, _ptr(prm), _ptr(out))
    if rc:
        raise RuntimeError(f"oracle_nonref failed ({rc})")
    return out
