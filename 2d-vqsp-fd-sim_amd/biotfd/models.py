"""Synthetic poroelastic models and sources in the reference's input formats.

The reference ships no model-preparation program: the meaning of each of the
24 material tables follows from how the time-loop kernels use them
(``stress_nn.F90:53-59``, ``inh_v_uv_int.F90:45-54``, ``inh_qx_int.F90:44-52``)
and is confirmed by ``ppml.F90:168`` and ``calcul_velocity.F90:16-36``, which
rebuild ``[[REL1_U, -REL3_U], [-REL3_U, REL2_QX]]`` as the inverse density
matrix and ``[[SIG_XX_1, -PRES_1], [-PRES_1, -PRES_3]]`` as the stiffness
matrix. With

    rho = (1-phi) rho_s + phi rho_f,   m = T rho_f / phi,   D = rho m - rho_f^2,
    b = eta / k,   alpha = 1 - Kdry/Ks,   M = 1 / (phi/Kf + (alpha-phi)/Ks),
    lam_c = Kdry - 2 mu / 3 + alpha^2 M,

the raw (unscaled) table entries are

    MU = mu,  SIG_XX_1 = SIG_ZZ_2 = lam_c + 2 mu,  SIG_XX_2 = SIG_ZZ_1 = lam_c,
    SIG_XX_3 = SIG_ZZ_3 = alpha M,  PRES_1 = PRES_2 = -alpha M,  PRES_3 = -M,
    REL1_U = m/D,  REL3_U = rho_f/D,  REL2_U = rho_f b/D,
    REL1_QX = rho_f/D,  REL2_QX = rho/D,  REL3_QX = rho b/D   (W/QZ identical),
    AUXIL_W = (lam_c - alpha^2 M) / (lam_c + 2 mu - alpha^2 M),
    AUXIL_QZ = 2 alpha mu / (lam_c + 2 mu - alpha^2 M).

The last two follow the commented model-preparation code kept at
``read_model_par.F90:78-87`` (``ALPHA_M = alpha M``, ``CAP_M = M``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from .formats import Job, MO_INDEX

RHO_S = 2650.0
RHO_F = 1000.0
K_S = 37e9
K_F = 2.25e9
ETA = 1e-3
TORT = 2.0

A4 = -1.0 / 24.0
B4 = 9.0 / 8.0


@dataclass(frozen=True)
class Rock:
    phi: float
    kdry: float
    mu: float
    perm: float
    rho_s: float = RHO_S
    rho_f: float = RHO_F
    ks: float = K_S
    kf: float = K_F
    eta: float = ETA
    tort: float = TORT

    def table_row(self) -> np.ndarray:
        """The 24 raw ``.MO`` entries of this rock (float64)."""
        return table_rows(np.array([self.phi]), np.array([self.kdry]), np.array([self.mu]),
                          np.array([self.perm]), self.rho_s, self.rho_f, self.ks, self.kf,
                          self.eta, self.tort)[0]

    def speeds(self):
        """(fast P, slow P, S) speeds in the no-friction limit, as calcul_velocity.F90."""
        return wave_speeds(self.table_row())


def table_rows(phi, kdry, mu, perm, rho_s=RHO_S, rho_f=RHO_F, ks=K_S, kf=K_F, eta=ETA,
               tort=TORT) -> np.ndarray:
    """Raw ``.MO`` rows ``[n][24]`` (float64) of n rocks given as arrays of
    porosity, dry bulk modulus, shear modulus and permeability."""
    phi, kdry, mu, perm = (np.asarray(v, np.float64) for v in (phi, kdry, mu, perm))
    rs, rf = rho_s, rho_f
    rho = (1.0 - phi) * rs + phi * rf
    m = tort * rf / phi
    d = rho * m - rf * rf
    b = eta / perm
    alpha = 1.0 - kdry / ks
    M = 1.0 / (phi / kf + (alpha - phi) / ks)
    lamc = kdry - 2.0 * mu / 3.0 + alpha * alpha * M
    l2mc = lamc + 2.0 * mu
    rows = np.zeros((phi.size, 24))
    v = {
        "MU": mu,
        "SIG_XX_1": l2mc, "SIG_XX_2": lamc, "SIG_XX_3": alpha * M,
        "SIG_ZZ_1": lamc, "SIG_ZZ_2": l2mc, "SIG_ZZ_3": alpha * M,
        "PRES_1": -alpha * M, "PRES_2": -alpha * M, "PRES_3": -M,
        "REL1_U": m / d, "REL1_W": m / d,
        "REL2_U": rf * b / d, "REL2_W": rf * b / d,
        "REL3_U": rf / d, "REL3_W": rf / d,
        "REL1_QX": rf / d, "REL1_QZ": rf / d,
        "REL2_QX": rho / d, "REL2_QZ": rho / d,
        "REL3_QX": rho * b / d, "REL3_QZ": rho * b / d,
        "AUXIL_W": (lamc - alpha * alpha * M) / (l2mc - alpha * alpha * M),
        "AUXIL_QZ": 2.0 * alpha * mu / (l2mc - alpha * alpha * M),
    }
    for k, val in v.items():
        rows[:, MO_INDEX[k]] = val
    return rows


def max_fast_speed(rows: np.ndarray) -> float:
    """Largest fast-P speed over table rows ``[n][24]`` (the expressions of
    :func:`wave_speeds`, vectorized)."""
    g = lambda k: rows[:, MO_INDEX[k]]
    a11, a12, a22 = g("REL1_U"), -g("REL3_U"), g("REL2_QX")
    b11, b12, b22 = g("SIG_XX_1"), -g("PRES_1"), -g("PRES_3")
    t00 = a11 * b11 + a12 * b12
    t01 = a11 * b12 + a12 * b22
    t10 = a12 * b11 + a22 * b12
    t11 = a12 * b12 + a22 * b22
    tr = t00 + t11
    delta = tr * tr + 4.0 * (t01 * t10 - t00 * t11)
    return float(np.sqrt((tr + np.sqrt(delta)) / 2.0).max())


def wave_speeds(row: np.ndarray):
    """Eigen-speeds of one raw table row (calcul_velocity.F90:16-36, float64)."""
    g = lambda k: row[MO_INDEX[k]]
    ainv = np.array([[g("REL1_U"), -g("REL3_U")], [-g("REL3_U"), g("REL2_QX")]])
    bm = np.array([[g("SIG_XX_1"), -g("PRES_1")], [-g("PRES_1"), -g("PRES_3")]])
    tt = ainv @ bm
    tr = tt[0, 0] + tt[1, 1]
    delta = tr * tr + 4.0 * (tt[0, 1] * tt[1, 0] - tt[0, 0] * tt[1, 1])
    vpf = math.sqrt((tr + math.sqrt(delta)) / 2.0)
    vps = math.sqrt(max((tr - math.sqrt(delta)) / 2.0, 0.0))
    vs = math.sqrt(g("REL1_U") * g("MU")) if g("MU") != 0 else vpf
    return vpf, vps, vs


def stable_dt(h: float, vmax: float, frac: float = 0.5) -> float:
    """``frac`` x the 2-D CFL limit of the 4th-order staggered stencil."""
    return frac * h / (vmax * math.sqrt(2.0) * (abs(A4) + abs(B4)))


def ricker(f0: float, dt_sample: float, t_end: float | None = None, t0: float | None = None):
    """Ricker wavelet samples for ``STF.DAT``: zero-phase at ``t0`` (default 1.2/f0)."""
    if t0 is None:
        t0 = 1.2 / f0
    if t_end is None:
        t_end = 2.5 / f0
    n = int(math.floor(t_end / dt_sample)) + 1
    t = np.arange(n) * dt_sample
    a = (math.pi * f0 * (t - t0)) ** 2
    return t, (1.0 - 2.0 * a) * np.exp(-a)


def layered_rocks(n_layers: int, rng: np.random.Generator) -> list:
    """Per-layer parameters drawn as in the layered half-space workload."""
    rocks = []
    for _ in range(n_layers):
        phi = rng.uniform(0.05, 0.35)
        kdry = rng.uniform(5e9, 20e9)
        mu = rng.uniform(3e9, 15e9)
        perm = 10.0 ** rng.uniform(-14.0, -11.0)
        rocks.append(Rock(phi, kdry, mu, perm))
    return rocks


def layer_index(mz: int, n_layers: int, rng: np.random.Generator) -> np.ndarray:
    """Row -> 0-based layer index for ``mz+1`` physical rows, seeded random depths."""
    rows = mz + 1
    if n_layers <= 1:
        return np.zeros(rows, dtype=np.int64)
    cuts = np.sort(rng.choice(np.arange(1, rows), size=n_layers - 1, replace=False))
    lay = np.zeros(rows, dtype=np.int64)
    for k, c in enumerate(cuts, start=1):
        lay[c:] = k
    return lay


def gaussian_field(shape, corr_cells: float, rng: np.random.Generator) -> np.ndarray:
    """Unit-variance Gaussian-correlated random field (spectral filter, float64)."""
    nz, nx = shape
    w = rng.standard_normal((nz, nx))
    kz = np.fft.fftfreq(nz)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    filt = np.exp(-((math.pi * corr_cells) ** 2) * (kx * kx + kz * kz) / 2.0)
    f = np.fft.irfft2(np.fft.rfft2(w) * filt, s=(nz, nx))
    f -= f.mean()
    s = f.std()
    return f / s if s > 0 else f


def receivers_row(n: int, x0_cell: int, dx_cells: int, row: int):
    """``n`` receivers on one grid row (same node for the U and W samples)."""
    return [(x0_cell + k * dx_cells, row, x0_cell + k * dx_cells, row) for k in range(n)]


def make_job(name: str, rocks, jm: np.ndarray, h: float, steps: int,
             f0: float, x_source: float, z_source: float, receivers=(), dt: float | None = None,
             **kw) -> Job:
    """Assemble a :class:`Job` whose run has exactly ``steps`` time steps;
    ``rocks`` is a sequence of :class:`Rock` or a raw table ``[JMNUM][24]``."""
    table = rocks if isinstance(rocks, np.ndarray) else np.stack([r.table_row() for r in rocks])
    if dt is None:
        vmax = max_fast_speed(table)
        dt = stable_dt(h, vmax)
    dt = float(np.float32(dt))
    # MT is chosen so that INT(MT/DT)+1 == steps in default-real arithmetic.
    mt = float(np.float32((steps - 0.5) * dt))
    t, v = ricker(f0, dt / 4.0)
    mz = jm.shape[0] - 1
    mx = jm.shape[1]
    job = Job(name=name, mx=mx, mz=mz, h=float(np.float32(h)), dt=dt, mt=mt, jm=jm,
              table=table, stf_t=t, stf_v=v, x_source=x_source, z_source=z_source,
              receivers=list(receivers), **kw)
    if job.mt2 != steps:
        raise AssertionError(f"MT2={job.mt2} != {steps}")
    return job


def homogeneous_job(name="HOMOG", mx=400, mz=400, h=1.0, steps=1500, tpml=20, f0=30.0,
                    n_rec=100, rec_dx=4, **kw) -> Job:
    """Homogeneous Biot model with a centred explosive source (the CPU-runnable case).

    Receivers sit 100 m above the source, every ``rec_dx`` cells across the model
    (100 receivers at 5 m would span 495 m of a 400 m model, so 4 m is used).
    """
    rock = Rock(phi=0.2, kdry=10e9, mu=9e9, perm=1e-12)
    jm = np.ones((mz + 1, mx), dtype=np.int32)
    xs, zs = mx * h / 2.0, mz * h / 2.0
    rrow = int(round((zs - 100.0) / h)) + 2
    x0 = max(1, (mx - (n_rec - 1) * rec_dx) // 2)
    recs = receivers_row(n_rec, x0, rec_dx, rrow)
    return make_job(name, [rock], jm, h, steps, f0, xs, zs, recs, tpml=tpml,
                    key_tlp=True, **kw)


def layered_job(name="LAYER", mx=4000, mz=2000, h=2.0, steps=5000, n_layers=10, seed=1,
                tpml=20, f0=25.0, n_rec=1000, field_std=0.0, corr_len=50.0, levels=0,
                **kw) -> Job:
    """Layered half-space with a surface source; optional correlated perturbation.

    With ``field_std > 0`` porosity, Kdry and mu are multiplied by
    ``1 + field_std * g`` for a Gaussian-correlated ``g`` (correlation length
    ``corr_len`` metres). ``levels`` quantizes ``g`` to that many values so the
    material table stays indexed (``levels**3 * n_layers`` entries at most);
    ``levels=None`` gives every cell its own material (JMNUM = cells).
    """
    rng = np.random.default_rng(seed)
    rocks = layered_rocks(n_layers, rng)
    lay = layer_index(mz, n_layers, rng)
    if field_std > 0.0 and levels is None:
        g = [gaussian_field((mz + 1, mx), corr_len / h, rng) for _ in range(3)]
        base = [np.array([getattr(r, k) for r in rocks])[lay][:, None] for k in ("phi", "kdry", "mu", "perm")]
        f = [1.0 + field_std * x for x in g]
        table = table_rows((base[0] * f[0]).ravel(), (base[1] * f[1]).ravel(),
                           (base[2] * f[2]).ravel(), np.broadcast_to(base[3], (mz + 1, mx)).ravel())
        del g, f
        jm = np.arange(1, (mz + 1) * mx + 1, dtype=np.int32).reshape(mz + 1, mx)
        vmax = max_fast_speed(table)
        dt = stable_dt(h, vmax)
        xs = float(rng.uniform(0.25, 0.75) * mx * h)
        n_rec = min(n_rec, mx)
        step = max(1, mx // n_rec)
        recs = receivers_row(n_rec, 1, step, 2)
        return make_job(name, table, jm, h, steps, f0, xs, 0.0, recs, dt=dt, tpml=tpml,
                        key_tlp=True, **kw)
    if field_std <= 0.0:
        jm = np.repeat((lay + 1)[:, None], mx, axis=1).astype(np.int32)
        table_rocks = rocks
    else:
        q = max(int(levels), 2)
        g = [gaussian_field((mz + 1, mx), corr_len / h, rng) for _ in range(3)]
        edges = np.linspace(-3.0, 3.0, q + 1)
        centers = 0.5 * (edges[:-1] + edges[1:])
        qi = [np.clip(np.digitize(x, edges[1:-1]), 0, q - 1) for x in g]
        code = ((lay[:, None] * q + qi[0]) * q + qi[1]) * q + qi[2]
        used, inv = np.unique(code, return_inverse=True)
        jm = (inv.reshape(code.shape) + 1).astype(np.int32)
        table_rocks = []
        for c in used:
            c = int(c)
            k2 = c % q; c //= q
            k1 = c % q; c //= q
            k0 = c % q; c //= q
            base = rocks[c]
            table_rocks.append(Rock(
                phi=base.phi * (1.0 + field_std * centers[k0]),
                kdry=base.kdry * (1.0 + field_std * centers[k1]),
                mu=base.mu * (1.0 + field_std * centers[k2]),
                perm=base.perm))
    vmax = max(r.speeds()[0] for r in table_rocks)
    dt = stable_dt(h, vmax)
    xs = float(rng.uniform(0.25, 0.75) * mx * h)
    n_rec = min(n_rec, mx)
    step = max(1, mx // n_rec)
    recs = receivers_row(n_rec, 1, step, 2)
    return make_job(name, table_rocks, jm, h, steps, f0, xs, 0.0, recs, dt=dt, tpml=tpml,
                    key_tlp=True, **kw)


def shot_set(n_shots=64, steps=3000, seed=1000, mx_range=(500, 3000), mz_range=(500, 1500),
             f0_range=(10.0, 40.0), h=2.0, tpml=20, n_rec=1000, n_layers=10) -> list:
    """The batched-shots workload: ``n_shots`` layered half-spaces (as
    :func:`layered_job`, shot s seeded with ``seed + s``) whose physical grid
    sizes are drawn uniformly from ``mx_range`` x ``mz_range``, each with a
    surface source at a random x and a Ricker frequency drawn from
    ``f0_range``, surface receivers, and ``steps`` time steps."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_shots):
        mx = int(rng.integers(mx_range[0], mx_range[1] + 1))
        mz = int(rng.integers(mz_range[0], mz_range[1] + 1))
        f0 = float(rng.uniform(*f0_range))
        out.append(layered_job(name=f"SHOT{s:03d}", mx=mx, mz=mz, h=h, steps=steps,
                               n_layers=n_layers, seed=seed + s, tpml=tpml, f0=f0, n_rec=n_rec))
    return out
