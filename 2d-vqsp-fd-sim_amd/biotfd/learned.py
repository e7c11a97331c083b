"""Learned absorbing boundary (``include/biotfd_learned.h``): the networks,
their training on bands recorded from wide-PML runs, and the acceptance check
against the wide-PML run.

A grid with TPML = 8 runs the reference's method with an 8-cell CPML
(``ppml.F90:159-306``). With a :class:`Network` attached, its strip cells are
replaced after every step by the networks' prediction (BF16 matrix-core
inference on the device). The networks are fitted on the device by the
library's FP32 training kernels to bands recorded from runs of the same kind of
models with a 40-cell PML, where the strip's cells are the first 8 cells of the
wide PML. Whether the result is good enough is decided per run by the
reflected boundary energy: the energy of the difference to the wide-PML run
over the physical cells, relative to the wide run's energy, summed over a few
time levels; the learned boundary is accepted where that stays <= 1e-3,
otherwise the plain PML is the fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from . import lib, models
from .formats import Job
from .lib import BiotError, Simulation, _ptr

S = 8                 # strip cells
CELLS = 12            # band cells across a boundary
F = 8                 # fields
BAND = F * CELLS      # 96
CIN = 4 * BAND        # 384
H = 64
TAP = 5
NW1 = TAP * CIN * H
NW2 = TAP * H * H
NSIDE = NW1 + 2 * NW2
SIDES = ("left", "right", "bottom")
ACCEPT = 1e-3         # reflected-energy ratio a learned boundary must stay within
HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "weights", "lb_strip8.npz")

_bound = False


def _lib():
    global _bound
    so = lib.load()
    if not _bound:
        p = C.c_void_p
        sig = {
            "bq_band_size": ([C.c_int, C.c_int], C.c_int),
            "bq_band_extract": ([p, p], C.c_int),
            "bq_lb_create": ([p, p, C.c_int, C.c_float, C.c_int, C.POINTER(p)], C.c_int),
            "bq_strip_copy": ([p, p], C.c_int),
            "bq_lb_trainer_capture_residual": ([p, p, p, C.c_int], C.c_int),
            "bq_lb_destroy": ([p], None),
            "bq_lb_last_error": ([], C.c_char_p),
            "bq_lb_attach": ([p, p], C.c_int),
            "bq_lb_detach": ([p], C.c_int),
            "bq_lb_infer": ([p, C.c_int, C.c_int, p, p], C.c_int),
            "bq_field_misfit": ([p, p, C.c_int, p], C.c_int),
            "bq_lb_trainer_create": ([C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(p)], C.c_int),
            "bq_lb_trainer_destroy": ([p], None),
            "bq_lb_trainer_last_error": ([p], C.c_char_p),
            "bq_lb_trainer_capture": ([p, p, C.c_int, C.c_int], C.c_int),
            "bq_lb_trainer_sync": ([p], C.c_int),
            "bq_lb_trainer_scales": ([p, C.c_int, p], C.c_int),
            "bq_lb_trainer_set_weights": ([p, p], C.c_int),
            "bq_lb_trainer_get_weights": ([p, p], C.c_int),
            "bq_lb_trainer_step": ([p, C.c_int, p, C.c_float, p], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(so, name)
            f.argtypes = args
            f.restype = res
        _bound = True
    return so


EXPORTS = ("bq_band_size", "bq_band_extract", "bq_lb_create", "bq_lb_destroy", "bq_lb_last_error",
           "bq_lb_attach", "bq_lb_detach", "bq_lb_infer", "bq_field_misfit", "bq_strip_copy",
           "bq_lb_trainer_capture_residual", "bq_lb_trainer_create",
           "bq_lb_trainer_destroy", "bq_lb_trainer_last_error", "bq_lb_trainer_capture",
           "bq_lb_trainer_sync", "bq_lb_trainer_scales", "bq_lb_trainer_set_weights",
           "bq_lb_trainer_get_weights", "bq_lb_trainer_step")


# ------------------------------------------------------------- the networks

def init_weights(seed: int = 0) -> np.ndarray:
    """He-initialised weights [3][NSIDE] (the last layer scaled by 0.1, so an
    untrained network changes the strip little)."""
    rng = np.random.default_rng(seed)
    out = np.zeros((3, NSIDE), np.float32)
    for s in range(3):
        w1 = rng.standard_normal(NW1) * np.sqrt(2.0 / (TAP * CIN))
        w2 = rng.standard_normal(NW2) * np.sqrt(2.0 / (TAP * H))
        w3 = rng.standard_normal(NW2) * np.sqrt(2.0 / (TAP * H)) * 0.1
        out[s] = np.concatenate([w1, w2, w3]).astype(np.float32)
    return out


def split_weights(w_side: np.ndarray):
    """One side's weights -> (W1 [5][384][64], W2 [5][64][64], W3 [5][64][64])."""
    w1 = w_side[:NW1].reshape(TAP, CIN, H)
    w2 = w_side[NW1:NW1 + NW2].reshape(TAP, H, H)
    w3 = w_side[NW1 + NW2:].reshape(TAP, H, H)
    return w1, w2, w3


def torch_forward(w_side: np.ndarray, scale_side: np.ndarray, bands: np.ndarray,
                  gain: float = 1.0) -> np.ndarray:
    """The network of one side in FP32 (torch, CPU): bands [4][L][96] -> the
    newest level's strip values plus gain * scale * output, [L][8][8] (field,
    cell). The reference the device inference is checked against."""
    import torch
    x, inv = _inputs(bands, scale_side)
    ws = [torch.from_numpy(np.ascontiguousarray(w.transpose(2, 1, 0))) for w in split_weights(w_side)]
    h = torch.from_numpy(x.T[None].copy())                       # [1][384][L]
    for k, w in enumerate(ws):
        h = torch.nn.functional.conv1d(h, w, padding=2)
        if k < 2:
            h = torch.nn.functional.leaky_relu(h, 0.1)
    d = h[0].T.numpy().reshape(-1, F, S)                          # [L][field][cell]
    prev = bands[3].reshape(-1, F, CELLS)[:, :, :S]
    return (prev + np.float32(gain) * scale_side[None, :, None] * d).astype(np.float32)


def _inputs(bands, scale_side):
    inv = (1.0 / np.asarray(scale_side, np.float32)).astype(np.float32)
    lv = np.asarray(bands, np.float32).reshape(4, -1, F, CELLS) * inv[None, None, :, None]
    x = lv.transpose(1, 0, 2, 3).reshape(-1, CIN)                 # [L][level*96 + f*12 + c]
    return np.ascontiguousarray(x, np.float32), inv


class Network:
    """Weights [3][NSIDE] and scales [3][8] of the three sides, the form
    (residual: correcting the 8-cell CPML's update; damping: the same
    correction limited to pulling each value toward zero; or direct) and the
    output gain; on the device as a ``bq_lb`` for :meth:`attach`."""

    def __init__(self, weights: np.ndarray, scales: np.ndarray, residual: bool = True,
                 gain: float = 1.0, inner: int = S, damping: bool = False,
                 profile: str = "cpml8"):
        self.weights = np.ascontiguousarray(weights, np.float32).reshape(3, NSIDE)
        self.scales = np.ascontiguousarray(scales, np.float32).reshape(3, F)
        self.residual = bool(residual)
        self.damping = bool(damping) and self.residual
        self.gain = float(gain)
        self.inner = int(inner)
        if profile not in PROFILES:
            raise ValueError(f"profile must be one of {PROFILES}")
        self.profile = profile
        self._h = None

    def strip_grid(self, job8: Job, wide: Simulation | None = None) -> Simulation:
        """A Simulation of job8 (TPML = 8) with the PML this network was
        trained on: the 8-cell CPML ("cpml8"), or the inner 8 nodes of the
        40-cell run's profile ("inner40", taken from wide, a TPML = 40
        Simulation of the same model)."""
        if self.profile == "cpml8":
            return Simulation(job8)
        if wide is None:
            raise ValueError("the inner40 profile needs the 40-cell run of the same model")
        return Simulation(job8, pml_arrays=inner_profile_pml(wide, job8.tpml))

    @property
    def form(self) -> int:
        """0 direct, 1 residual, 2 damping (bq_lb_create's form)."""
        return 2 if self.damping else int(self.residual)

    @classmethod
    def load(cls, path: str = WEIGHTS, gain: float | None = None) -> "Network":
        z = np.load(path, allow_pickle=False)
        g = float(z["gain"]) if gain is None and "gain" in z else (1.0 if gain is None else gain)
        return cls(z["weights"], z["scales"], bool(z["residual"]) if "residual" in z else True, g,
                   int(z["inner"]) if "inner" in z else S,
                   bool(z["damping"]) if "damping" in z else False,
                   str(z["profile"]) if "profile" in z else "cpml8")

    def with_gain(self, gain: float, damping: bool | None = None) -> "Network":
        return Network(self.weights, self.scales, self.residual, gain, self.inner,
                       self.damping if damping is None else damping, self.profile)

    def save(self, path: str, **info):
        os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
        np.savez_compressed(path, weights=self.weights, scales=self.scales,
                            residual=np.asarray(self.residual), damping=np.asarray(self.damping),
                            gain=np.asarray(self.gain),
                            inner=np.asarray(self.inner), profile=np.asarray(self.profile),
                            **{k: np.asarray(v) for k, v in info.items()})

    def device(self):
        if self._h is None:
            h = C.c_void_p()
            rc = _lib().bq_lb_create(_ptr(self.weights), _ptr(self.scales), self.form,
                                     float(self.gain), self.inner, C.byref(h))
            if rc:
                raise BiotError(f"bq_lb_create failed ({rc}): {_lib().bq_lb_last_error().decode()}")
            self._h = h
        return self._h

    def attach(self, sim: Simulation):
        sim._chk(_lib().bq_lb_attach(sim.h, self.device()))

    def infer(self, side: int, bands: np.ndarray) -> np.ndarray:
        """Device inference (BF16 MFMA) on bands [4][L][96] -> [L][8][8]."""
        b = np.ascontiguousarray(bands, np.float32)
        L = b.shape[1]
        out = np.zeros((L, F, S), np.float32)
        rc = _lib().bq_lb_infer(self.device(), int(side), int(L), _ptr(b), _ptr(out))
        if rc:
            raise BiotError(f"bq_lb_infer failed ({rc})")
        return out

    def close(self):
        if self._h is not None:
            _lib().bq_lb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PROFILES = ("cpml8", "inner40")


def inner_profile_pml(wide: Simulation, tpml: int = S) -> list:
    """The 24 PML coefficient arrays (bq_upload_pml order) of a grid with TPML
    = tpml cut from those of wide, a run of the same model with a wider PML:
    every strip cell gets the coefficients the wide run's cell at the same
    place has, so a truncated grid whose strip holds the wide run's values
    advances every cell at least 3 from its outer edge exactly as the wide run
    does. Layouts: x arrays [row][node] over the rows, z arrays [column][node]
    over the columns (nodes 0..T or 0..T+1)."""
    tw = wide.job.tpml
    if tw < tpml:
        raise ValueError("the wide run's PML must be at least as wide")
    nzw, nxw = wide.nz, wide.nx
    nz, nx = wide.job.mzp + tpml + 1, wide.job.mx + 2 * tpml
    out = []
    for k in range(24):
        a = wide.pml(k)
        n = a.size
        if n == (tw + 1) * nzw:
            out.append(np.ascontiguousarray(a.reshape(nzw, tw + 1)[:nz, :tpml + 1]))
        elif n == (tw + 2) * nzw:
            out.append(np.ascontiguousarray(a.reshape(nzw, tw + 2)[:nz, :tpml + 2]))
        elif n == (tw + 1) * nxw:
            out.append(np.ascontiguousarray(a.reshape(nxw, tw + 1)[tw - tpml:tw - tpml + nx, :tpml + 1]))
        else:
            out.append(np.ascontiguousarray(a.reshape(nxw, tw + 2)[tw - tpml:tw - tpml + nx, :tpml + 2]))
    return out


def band_host(fields: dict, mx: int, mzp: int, tpml: int) -> np.ndarray:
    """The band of one time level from host copies of the 8 fields ([NZc][NXc]
    each) of a grid with TPML >= 8, as bq_band_extract lays it out."""
    pmll = 1 - tpml
    names = ("U", "W", "QX", "QZ", "TXX", "TZZ", "TXZ", "P")
    lz, lx = mzp + 1, mx + 2 * S
    out = []
    c = np.arange(CELLS)
    for side in range(3):
        if side == 0:
            ii, ll = (1 - S + c)[None, :].repeat(lz, 0), np.arange(lz)[:, None].repeat(CELLS, 1)
        elif side == 1:
            ii, ll = (mx + S - c)[None, :].repeat(lz, 0), np.arange(lz)[:, None].repeat(CELLS, 1)
        else:
            ii = (1 - S + np.arange(lx))[:, None].repeat(CELLS, 1)
            ll = (mzp + S - c)[None, :].repeat(lx, 0)
        vals = np.stack([fields[n][ll, ii - pmll] for n in names], axis=1)   # [L][8][12]
        out.append(vals.reshape(-1))
    return np.concatenate(out).astype(np.float32)


def strip_copy(dst: Simulation, src: Simulation):
    """The 8-cell strip of dst's current fields from src's (bq_strip_copy)."""
    dst._chk(_lib().bq_strip_copy(dst.h, src.h))


def field_misfit(a: Simulation, b: Simulation, field: str) -> np.ndarray:
    """(sum (a-b)^2, sum b^2) of one field over the common physical cells."""
    out = np.zeros(2, np.float64)
    a._chk(_lib().bq_field_misfit(a.h, b.h, lib.FIELDS.index(field), _ptr(out)))
    return out


# ------------------------------------------------------------- the trainer

class Trainer:
    """Band store and FP32 training kernels on the device (bq_lb_trainer)."""

    def __init__(self, mx: int, mzp: int, max_groups: int, batch: int = 8, residual: bool = True,
                 inner: int = S):
        h = C.c_void_p()
        rc = _lib().bq_lb_trainer_create(mx, mzp, max_groups, batch, int(residual), int(inner),
                                         C.byref(h))
        if rc:
            raise BiotError(f"bq_lb_trainer_create failed ({rc}): "
                            f"{_lib().bq_lb_trainer_last_error(None).decode()}")
        self.t, self.mx, self.mzp, self.max_groups, self.batch = h, mx, mzp, max_groups, batch
        self.residual = bool(residual)
        self.inner = int(inner)
        self.ngroups = 0
        self.scales = np.ones((3, F), np.float32)

    def _chk(self, rc):
        if rc:
            raise BiotError(f"trainer error {rc}: {_lib().bq_lb_trainer_last_error(self.t).decode()}")

    def capture(self, sim: Simulation, group: int, level: int):
        self._chk(_lib().bq_lb_trainer_capture(self.t, sim.h, group, level))

    def capture_residual(self, wide: Simulation, trunc: Simulation, group: int):
        self._chk(_lib().bq_lb_trainer_capture_residual(self.t, wide.h, trunc.h, group))

    def compute_scales(self, ngroups: int) -> np.ndarray:
        out = np.zeros((3, F), np.float32)
        self._chk(_lib().bq_lb_trainer_scales(self.t, ngroups, _ptr(out)))
        self.scales = out
        return out

    def set_weights(self, w: np.ndarray):
        w = np.ascontiguousarray(w, np.float32)
        self._chk(_lib().bq_lb_trainer_set_weights(self.t, _ptr(w)))

    def get_weights(self) -> np.ndarray:
        w = np.zeros((3, NSIDE), np.float32)
        self._chk(_lib().bq_lb_trainer_get_weights(self.t, _ptr(w)))
        return w

    def step(self, side: int, groups, lr: float) -> float:
        g = np.ascontiguousarray(groups, np.int32)
        loss = np.zeros(1, np.float64)
        self._chk(_lib().bq_lb_trainer_step(self.t, int(side), _ptr(g), float(lr), _ptr(loss)))
        return float(loss[0])

    def close(self):
        if getattr(self, "t", None):
            _lib().bq_lb_trainer_destroy(self.t)
            self.t = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def training_job(seed: int, steps: int, tpml: int = 40, mx: int = 1000, mz: int = 1000):
    """A training model: a seeded layered half-space (as the layered workload)
    of mx x mz cells, dx = 2 m, with a source at a random position (x and
    depth) and a Ricker frequency drawn from 10..40 Hz; no receivers."""
    rng = np.random.default_rng(50_000 + seed)
    f0 = float(rng.uniform(10.0, 40.0))
    job = models.layered_job(name=f"TRAIN{seed:04d}", mx=mx, mz=mz, h=2.0, steps=steps,
                             n_layers=10, seed=seed, tpml=tpml, f0=f0, n_rec=1)
    job.receivers = []
    job.x_source = float(rng.uniform(0.05, 0.95) * mx * job.h)
    job.z_source = float(rng.uniform(0.0, 0.9) * mz * job.h)
    return job


def record(trainer: Trainer, n_models: int, steps: int, first: int, every: int,
           per_batch: int = 50, log=print, profile: str = "cpml8") -> int:
    """Runs n_models wide-PML (TPML = 40) training models, per_batch at a time
    as one batch, and records one group per model at t = first, first + every,
    ... < steps - 1: levels t-3 .. t of the wide run as input and the target
    level. Residual trainer: every model also runs on an 8-cell grid whose
    strip is copied from the wide run after every step, so that it stays on
    the wide run's trajectory; level 4 is the wide band at t+1 minus that
    grid's band after its own step t+1 (before the copy). Direct trainer:
    level 4 is the wide band at t+1. profile: the truncated grid's PML, the
    8-cell CPML or the inner 8 nodes of the 40-cell one (inner_profile_pml).
    Returns the number of groups."""
    times = list(range(first, steps - 1, every))
    # steps at which each level is taken: t-3, t-2, t-1, t (inputs), t+1
    when = {}
    for gi, t in enumerate(times):
        for lev in range(5):
            when.setdefault(t - 3 + lev, []).append((gi, lev))
    g0 = 0
    for b0 in range(0, n_models, per_batch):
        seeds = list(range(b0, min(n_models, b0 + per_batch)))
        if g0 + len(seeds) * len(times) > trainer.max_groups:
            break
        wide = [Simulation(training_job(s, steps, 40)) for s in seeds]
        bw = lib.Batch(wide)
        trunc = bt = None
        if trainer.residual:
            if profile == "inner40":
                trunc = [Simulation(training_job(s, steps, 8), pml_arrays=inner_profile_pml(w))
                         for s, w in zip(seeds, wide)]
            else:
                trunc = [Simulation(training_job(s, steps, 8)) for s in seeds]
            bt = lib.Batch(trunc)
        t0 = time.perf_counter()
        last = max(when)
        done = 0
        step = 0
        while step < last:
            # without a truncated grid, run straight to the next recorded level
            nxt = step + 1 if trainer.residual else min(k for k in when if k > step)
            bw.run(nxt - step)
            if bt is not None:
                bt.run(nxt - step)
            step = nxt
            for gi, lev in when.get(step, []):
                for k in range(len(seeds)):
                    grp = g0 + gi * len(seeds) + k
                    if lev == 4 and trainer.residual:
                        trainer.capture_residual(wide[k], trunc[k], grp)
                    else:
                        trainer.capture(wide[k], grp, lev)
            if trunc is not None:
                for k in range(len(seeds)):
                    strip_copy(trunc[k], wide[k])
        _lib().bq_lb_trainer_sync(trainer.t)
        g0 += len(seeds) * len(times)
        log(f"[lb] recorded models {seeds[0]}..{seeds[-1]}: {g0} groups "
            f"({time.perf_counter() - t0:.1f} s)")
        for b in (bw, bt):
            if b is not None:
                b.close()
        for s in wide + (trunc or []):
            s.close()
    trainer.ngroups = g0
    return g0


def train(trainer: Trainer, weights: np.ndarray, iters: int, lr: float, val_frac: float = 0.1,
          seed: int = 0, log=print) -> dict:
    """Adam on every side, iters minibatches each, learning rate lr with a
    cosine decay; groups split into training and validation sets. Returns the
    loss history."""
    rng = np.random.default_rng(seed)
    n = trainer.ngroups
    perm = rng.permutation(n)
    nval = max(trainer.batch, int(n * val_frac))
    val, tr = perm[:nval], perm[nval:]
    # the validation loss of a network that predicts nothing (zero output):
    # the variance the training has to explain
    trainer.set_weights(np.zeros_like(weights))
    base = [float(np.mean([trainer.step(side, val[k:k + trainer.batch], 0.0)
                           for k in range(0, len(val) - trainer.batch + 1, trainer.batch)]))
            for side in range(3)]
    log("[lb] zero-output validation loss: " + ", ".join(f"{SIDES[s]} {base[s]:.4e}" for s in range(3)))
    trainer.set_weights(weights)
    hist = {"train": [[] for _ in range(3)], "val": [[] for _ in range(3)], "zero_val": base}
    for side in range(3):
        for it in range(iters):
            a = lr * 0.5 * (1.0 + np.cos(np.pi * it / iters))
            loss = trainer.step(side, rng.choice(tr, trainer.batch, replace=False), a)
            hist["train"][side].append(loss)
            if (it + 1) % max(1, iters // 10) == 0 or it == iters - 1:
                vl = np.mean([trainer.step(side, val[k:k + trainer.batch], 0.0)
                              for k in range(0, len(val) - trainer.batch + 1, trainer.batch)])
                hist["val"][side].append(float(vl))
                log(f"[lb] {SIDES[side]} iter {it + 1}: train {loss:.4e} val {vl:.4e}")
    return hist


# ------------------------------------------------------------- acceptance

def reflected_energy(run: Simulation, wide: Simulation) -> np.ndarray:
    """(sum over U, W of (run - wide)^2, sum of wide^2) over the physical
    cells at the current time level."""
    return field_misfit(run, wide, "U") + field_misfit(run, wide, "W")
