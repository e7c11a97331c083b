"""Reference input/output file formats of the Biot velocity-stress simulator.

Every format here is the one the reference's Fortran reads or writes, so a job
directory produced by :func:`write_job` runs unchanged under the reference
driver (``main.F90``) and under the drop-in time loop of this package.

Input side (read by the reference):

* ``HF_2DFD_VS_PORO`` - first line is the job name (``main.F90:52-54``).
* ``<JOB>.IN`` - namelists ``CONTROLDATA``, ``CONTROLDATA1``, ``NAMES_SR_REC``,
  ``KEYS``, ``SOURCE``, ``NONREF``, ``TXT``, ``REC`` in that file order
  (``read_open.F90:112-218``).
* ``<JOB>.INM`` - namelists ``NAMES``, ``CONTROLDATA2`` (``read_open.F90:117,144``).
* ``<JOB>.JMH`` - unformatted sequential, one int32 record of ``MX`` material
  indices per physical row (``read_model.F90:57-77``).
* ``<JOB>.MO`` - unformatted sequential: record 1 ``JMNUM`` (int32), record 2
  the 24 material tables interleaved per material (``read_model_par.F90:37-60``).
* ``STF.DAT`` - list-directed ``time value`` pairs (``open_sr_file.F90:62-76``).
* ``REC.DAT`` - list-directed ``name irecx lrecx irecz lrecz`` per receiver
  (``read_open.F90:234-236``).

Output side (written by the reference):

* ``REC_UW_<j>.asc`` / ``REC_Q_<j>.asc`` - ``(5(1X,E13.6))`` per step
  (``store_seismograms.F90:140-153``).
* ``REC_STRESS_<j>.asc`` - ``(3(1X,E13.6))`` per step
  (``store_seismograms_pressure.F90:75-80``).
* ``SNAP.Vnnnnnn`` / ``SNAP.Qnnnnnn`` - one unformatted record holding
  ``UM, WM`` (``QX, QZ``) over the whole grid (``store_snapshots.F90:173-190``).

Unformatted sequential records use 4-byte little-endian length markers before
and after each record (the layout of the flang and gfortran runtimes).
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass, field
from typing import Iterable, List, Optional, Sequence

import numpy as np

# Order of the 24 material tables inside a .MO record (read_model_par.F90:52-60).
MO_FIELDS = (
    "MU", "SIG_XX_1", "SIG_XX_2", "SIG_XX_3",
    "SIG_ZZ_1", "SIG_ZZ_2", "SIG_ZZ_3",
    "PRES_1", "PRES_2", "PRES_3",
    "REL1_U", "REL1_W", "REL2_U",
    "REL2_W", "REL3_U", "REL3_W",
    "REL1_QX", "REL1_QZ", "REL2_QX",
    "REL2_QZ", "REL3_QX", "REL3_QZ",
    "AUXIL_W", "AUXIL_QZ",
)
MO_INDEX = {name: k for k, name in enumerate(MO_FIELDS)}


# ----------------------------------------------------------------------------
# Unformatted sequential records
# ----------------------------------------------------------------------------

def write_records(path: str, records: Iterable[bytes]) -> None:
    """Write Fortran unformatted sequential records with 4-byte markers."""
    with open(path, "wb") as f:
        for rec in records:
            n = len(rec)
            if n >= 2 ** 31:
                raise ValueError("record too long for 4-byte markers")
            f.write(struct.pack("<i", n))
            f.write(rec)
            f.write(struct.pack("<i", n))


def read_records(path: str) -> List[bytes]:
    """Read every record of a Fortran unformatted sequential file."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        pos += 4
        out.append(data[pos:pos + n])
        pos += n
        (m,) = struct.unpack_from("<i", data, pos)
        if m != n:
            raise ValueError(f"{path}: record marker mismatch ({n} vs {m})")
        pos += 4
    return out


# ----------------------------------------------------------------------------
# Namelists
# ----------------------------------------------------------------------------

def _fmt_value(v) -> str:
    if isinstance(v, bool):
        return ".TRUE." if v else ".FALSE."
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return repr(float(v)).replace("e", "E")
    if isinstance(v, str):
        return "'" + v.replace("'", "''") + "'"
    raise TypeError(f"unsupported namelist value {v!r}")


def format_namelist(group: str, values: dict) -> str:
    body = ",\n".join(f"  {k} = {_fmt_value(v)}" for k, v in values.items())
    return f"&{group}\n{body}\n/\n"


# ----------------------------------------------------------------------------
# Job description
# ----------------------------------------------------------------------------

@dataclass
class Job:
    """Everything the reference reads for one run (names follow read_open.F90).

    ``jm`` is the physical material-index grid, shape ``(MZ+1, MX)`` with the
    first row at the free surface, values 1-based into the material table.
    ``table`` is the raw (unscaled) material table, shape ``(JMNUM, 24)`` in the
    order of :data:`MO_FIELDS`.
    """

    name: str
    mx: int
    mz: int
    h: float
    dt: float
    mt: float                      # record length in seconds (CONTROLDATA1 MT)
    jm: np.ndarray
    table: np.ndarray
    stf_t: np.ndarray
    stf_v: np.ndarray
    x_source: float
    z_source: float
    receivers: Sequence[tuple] = ()   # (irecx, lrecx, irecz, lrecz)
    tpml: int = 20
    stress_imaging: bool = True
    # the reference's compile-time options, as bits: 1 = AFDA_ZZ
    # (strain_uw_zz_0.F90:52), 2 = AFDA_XZ (strain_uw_xz_1.F90:83), 4 = SECOND
    # (mod_auxil.F90:15-21: A = 0, B = 1, second-order stencils)
    variant: int = 0
    r: float = 0.001
    tau: float = 1.5
    pow: float = 2.0
    sf: float = 0.0
    wc: float = 0.0
    ipas2: int = 1000
    key_tlv: bool = True
    key_snv: bool = False
    key_tlq: bool = False
    key_tlp: bool = False
    key_sour: int = 0
    qr: float = 0.0
    xbmin: float = 0.0
    ktbo: int = 0
    ktle: int = 0
    ktri: int = 0
    omg: float = 0.0
    # NONREF namelist parameters of the Higdon / optimal-absorption boundary
    # coefficients (read_open.F90:173-179 defaults)
    wb: float = 0.4
    thpple: float = 0.19757
    thssle: float = 0.70511
    text: str = "biotfd synthetic model"
    extra: dict = field(default_factory=dict)

    # -- derived quantities, as the reference derives them -------------------
    @property
    def mzp(self) -> int:
        """MZ after read_open.F90:149-151 (plus 2 image rows when SI)."""
        return self.mz + 2 if self.stress_imaging else self.mz

    @property
    def mt2(self) -> int:
        """Step count, MT2 = INT(MT/DT)+1 in default-real arithmetic (read_open.F90:153)."""
        return int(np.float32(self.mt) / np.float32(self.dt)) + 1

    @property
    def lin_x(self) -> int:
        xs = np.float32(abs(self.x_source))
        return _nint((xs - np.float32(self.xbmin)) / np.float32(self.h))

    @property
    def lin_z(self) -> int:
        zs = np.float32(abs(self.z_source))
        if self.stress_imaging:
            zs = np.float32(zs + np.float32(2.0) * np.float32(self.h))
        return _nint(zs / np.float32(self.h))


def _nint(x) -> int:
    """Fortran NINT: round half away from zero."""
    x = float(x)
    return int(np.floor(x + 0.5)) if x >= 0 else -int(np.floor(-x + 0.5))


def write_job(job: Job, directory: str, real_dtype=np.float32) -> None:
    """Write every input file of ``job`` into ``directory``.

    ``real_dtype`` is the kind of the ``.MO`` reals: float32 for the reference's
    default ``PP=SP`` build, float64 for an FP64 rebuild.
    """
    os.makedirs(directory, exist_ok=True)
    p = lambda n: os.path.join(directory, n)
    with open(p("HF_2DFD_VS_PORO"), "w") as f:
        f.write(job.name + "\n")

    with open(p(job.name + ".IN"), "w") as f:
        f.write(format_namelist("CONTROLDATA", dict(
            TPML=job.tpml, PROCX=1, PROCZ=1, STRESS_IMAGING=job.stress_imaging)))
        f.write(format_namelist("CONTROLDATA1", dict(
            MT=job.mt, R=job.r, TAU=job.tau, POW=job.pow, SF=job.sf, WC=job.wc,
            TAU_EPS=0.0, IPAS2=job.ipas2)))
        f.write(format_namelist("NAMES_SR_REC", dict(
            SR_FILE_NAME="STF.DAT", REC_FILE_NAME="REC.DAT")))
        f.write(format_namelist("KEYS", dict(
            KEY_TLV=job.key_tlv, KEY_SNV=job.key_snv, KEY_TLQ=job.key_tlq,
            KEY_TLP=job.key_tlp)))
        f.write(format_namelist("SOURCE", dict(
            X_SOURCE=job.x_source, Z_SOURCE=job.z_source, KEY_SOUR=job.key_sour,
            QR=job.qr)))
        f.write(format_namelist("NONREF", dict(
            OMG=job.omg, WB=job.wb, THPPLE=job.thpple, THSSLE=job.thssle, KTLE=job.ktle,
            KTRI=job.ktri, KTBO=job.ktbo)))
        f.write(format_namelist("TXT", dict(TEXT=job.text[:40])))
        f.write(format_namelist("REC", dict(MR=len(job.receivers))))

    with open(p(job.name + ".INM"), "w") as f:
        f.write(format_namelist("NAMES", dict(
            MO_FILE_NAME=job.name + ".MO", JMH_FILE_NAME=job.name + ".JMH")))
        f.write(format_namelist("CONTROLDATA2", dict(
            MX=job.mx, MZ=job.mz, H=job.h, DT_VS=job.dt, XBMIN=job.xbmin)))

    write_jmh(p(job.name + ".JMH"), job.jm)
    write_mo(p(job.name + ".MO"), job.table, real_dtype)
    write_stf(p("STF.DAT"), job.stf_t, job.stf_v)
    write_rec(p("REC.DAT"), job.receivers)


def write_jmh(path: str, jm: np.ndarray) -> None:
    jm = np.ascontiguousarray(jm, dtype="<i4")
    write_records(path, (row.tobytes() for row in jm))


def read_jmh(path: str, mx: int) -> np.ndarray:
    recs = read_records(path)
    return np.stack([np.frombuffer(r, dtype="<i4", count=mx) for r in recs])


def write_mo(path: str, table: np.ndarray, real_dtype=np.float32) -> None:
    table = np.asarray(table)
    if table.ndim != 2 or table.shape[1] != 24:
        raise ValueError("material table must be (JMNUM, 24)")
    rec2 = np.ascontiguousarray(table, dtype=np.dtype(real_dtype).newbyteorder("<"))
    write_records(path, [struct.pack("<i", table.shape[0]), rec2.tobytes()])


def read_mo(path: str, real_dtype=np.float32) -> np.ndarray:
    r1, r2 = read_records(path)[:2]
    (n,) = struct.unpack("<i", r1)
    n = abs(n)
    return np.frombuffer(r2, dtype=np.dtype(real_dtype).newbyteorder("<")).reshape(n, 24)


def write_stf(path: str, t: np.ndarray, v: np.ndarray) -> None:
    """Write source samples; values are rounded to float32 first and printed
    with 9 significant digits so that a float32 list-directed read recovers
    exactly the same numbers."""
    t32 = np.asarray(t, dtype=np.float32)
    v32 = np.asarray(v, dtype=np.float32)
    with open(path, "w") as f:
        for a, b in zip(t32.astype(np.float64), v32.astype(np.float64)):
            f.write(f"{a:.8E} {b:.8E}\n")


def read_stf(path: str):
    d = np.loadtxt(path, ndmin=2)
    return d[:, 0], d[:, 1]


def write_rec(path: str, receivers: Sequence[tuple]) -> None:
    with open(path, "w") as f:
        for j, (ix, lx, iz, lz) in enumerate(receivers, start=1):
            f.write(f"R{j:05d} {int(ix)} {int(lx)} {int(iz)} {int(lz)}\n")


# ----------------------------------------------------------------------------
# Output files
# ----------------------------------------------------------------------------

def read_snapshot(path: str, nx: int, nz: int, real_dtype=np.float32):
    """Return the two fields of a ``SNAP.*`` file as ``[NZc][NXc]`` arrays."""
    (rec,) = read_records(path)[:1]
    a = np.frombuffer(rec, dtype=np.dtype(real_dtype).newbyteorder("<"))
    if a.size != 2 * nx * nz:
        raise ValueError(f"{path}: expected {2 * nx * nz} reals, got {a.size}")
    return a[: nx * nz].reshape(nz, nx), a[nx * nz:].reshape(nz, nx)


def write_snapshot(path: str, a: np.ndarray, b: np.ndarray, real_dtype=np.float32) -> None:
    rec = (np.ascontiguousarray(a, dtype=real_dtype).tobytes()
           + np.ascontiguousarray(b, dtype=real_dtype).tobytes())
    write_records(path, [rec])


def read_rec_asc(path: str) -> np.ndarray:
    return np.loadtxt(path, ndmin=2)
