"""Small seeded jobs shared by the tests (sizes the CPU oracle runs in seconds)."""
import numpy as np

from biotfd import models
from biotfd.formats import Job


def tiny_layered(steps=120, tpml=8, seed=3, mx=48, mz=36, f0=60.0) -> Job:
    """Three random layers, surface source: waves reach the side PML and the surface."""
    return models.layered_job(name="TINY", mx=mx, mz=mz, h=2.0, steps=steps, n_layers=3, seed=seed,
                              tpml=tpml, f0=f0, n_rec=10)


def tiny_homog(steps=120, tpml=6, mx=40, mz=30, f0=80.0) -> Job:
    """Homogeneous model, centred explosive source, receivers on a row above it."""
    rock = models.Rock(phi=0.2, kdry=10e9, mu=9e9, perm=1e-12)
    jm = np.ones((mz + 1, mx), dtype=np.int32)
    recs = models.receivers_row(8, 4, 4, 8)
    return models.make_job("HOMOG", [rock], jm, 1.0, steps, f0, mx / 2.0, mz / 2.0, recs,
                           tpml=tpml, key_tlp=True)


def tiny_edges(steps=150, tpml=5, mx=36, mz=24, f0=90.0, seed=7) -> Job:
    """Source in the left PML strip near the bottom row (LIN_Z = MZ'-2), so the
    bottom PML, the last row and the left strip all see large amplitudes;
    receivers in the PML, at the surface and out of the grid."""
    rng = np.random.default_rng(seed)
    rocks = models.layered_rocks(2, rng)
    jm = np.ones((mz + 1, mx), dtype=np.int32)
    jm[mz // 2:, :] = 2
    jm[:, mx // 2:] = np.where(jm[:, mx // 2:] == 1, 2, 1)
    h = 1.5
    xs, zs = 0.0, (mz + 2 - 2 - 2) * h   # LIN_X = 0 (left PML), LIN_Z = MZ' - 2
    recs = [(-tpml + 2, 3, -tpml + 2, 3), (0, 2, 0, 2), (mx + tpml - 1, mz + 2 + tpml - 1, mx, mz),
            (5, 1, 5, 1), (10, 0, 10, 0), (mx + tpml + 4, 3, 3, mz + tpml + 9)]
    return models.make_job("EDGES", rocks, jm, h, steps, f0, xs, zs, recs, tpml=tpml, key_tlp=True)


def tiny_layered_nosi(steps=120, variant=0, **kw) -> Job:
    """Explicit free surface (no stress imaging, MZ' = MZ): the surface rows 0
    and 1 use their own one-sided stencils. Source 3 cells deep (rows 0 and 1
    cannot hold it), receivers on rows 0 and 1. variant selects the reference's
    AFDA_ZZ / AFDA_XZ / SECOND build options."""
    job = tiny_layered(steps=steps, **kw)
    job.stress_imaging = False
    job.variant = variant
    job.z_source = 3.0 * job.h
    job.receivers = [(r[0], k % 2, r[2], k % 2) for k, r in enumerate(job.receivers)]
    return job


def tiny_edges_nosi(steps=150, **kw) -> Job:
    """The edge case without stress imaging: source in the left PML strip on
    row MZ'-2 = MZ-2."""
    job = tiny_edges(steps=steps, **kw)
    job.stress_imaging = False
    job.z_source = (job.mz - 2) * job.h
    return job


ALL = {"layered": tiny_layered, "homog": tiny_homog, "edges": tiny_edges,
       "layered_nosi": tiny_layered_nosi, "edges_nosi": tiny_edges_nosi,
       "layered_nosi_afda": lambda: tiny_layered_nosi(variant=3),
       "layered_second": lambda: _second(tiny_layered()),
       "layered_nosi_second": lambda: tiny_layered_nosi(variant=4)}


def _second(job: Job) -> Job:
    """The reference's second-order build (-DSECOND: A = 0, B = 1)."""
    job.variant = 4
    return job


def medium_layered(steps=200, tpml=10, mx=200, mz=120, seed=5, f0=40.0) -> Job:
    """Large enough that the interior fast tiles (64 columns wide) are used."""
    return models.layered_job(name="MEDIUM", mx=mx, mz=mz, h=2.0, steps=steps, n_layers=4, seed=seed,
                              tpml=tpml, f0=f0, n_rec=40)


def medium_deep(steps=200, tpml=12, mx=230, mz=140, f0=45.0) -> Job:
    """Correlated random model, source deep inside a fast tile."""
    job = models.layered_job(name="DEEP", mx=mx, mz=mz, h=2.0, steps=steps, n_layers=3, seed=9,
                             tpml=tpml, f0=f0, n_rec=30, field_std=0.08, corr_len=20.0, levels=6)
    job.x_source = 151.0 * 2.0
    job.z_source = 70.0 * 2.0
    return job


GPU_EXTRA = {"medium": medium_layered, "deep": medium_deep}


def wide_deep(steps=160, tpml=9, mx=1100, mz=170, f0=30.0, stress_imaging=True) -> Job:
    """Wide enough for two fused tiles of 512 columns side by side and two rows
    of 128-row blocks; correlated random model. The source cell (535, 132) is
    the last column and row of the first tile, so the tiles right of and below
    it recompute it in their halos. With TPML = 9 the tiles start at column 24."""
    job = models.layered_job(name="WIDE", mx=mx, mz=mz, h=2.0, steps=steps, n_layers=4, seed=13,
                             tpml=tpml, f0=f0, n_rec=50, field_std=0.08, corr_len=16.0, levels=5)
    job.stress_imaging = stress_imaging
    job.x_source = 535.0 * 2.0
    job.z_source = (132.0 - (2.0 if stress_imaging else 0.0)) * 2.0
    if not stress_imaging:
        job.receivers = [(r[0], 0, r[2], 0) for r in job.receivers]
    return job


GPU_EXTRA["wide"] = wide_deep
GPU_EXTRA["wide_nosi"] = lambda: wide_deep(stress_imaging=False)


def narrow_tiles(steps=140, tpml=10, mx=950, mz=150, f0=35.0, stress_imaging=True) -> Job:
    """Interior wide enough for one full fused tile (512 columns), a column of
    half tiles (256) and one of quarter tiles (128), then a masked run of 30
    columns: with TPML = 10 the tiles cover [23, 535), [535, 791), [791, 919)
    and the runs [919, 949). The source cell (790, 100) is the last column of
    the half tiles."""
    job = models.layered_job(name="NARROW", mx=mx, mz=mz, h=2.0, steps=steps, n_layers=4, seed=21,
                             tpml=tpml, f0=f0, n_rec=60, field_std=0.08, corr_len=16.0, levels=5)
    job.stress_imaging = stress_imaging
    job.x_source = 790.0 * 2.0
    job.z_source = (100.0 - (2.0 if stress_imaging else 0.0)) * 2.0
    if not stress_imaging:
        job.receivers = [(r[0], 0, r[2], 0) for r in job.receivers]
    return job


GPU_EXTRA["narrow"] = narrow_tiles
GPU_EXTRA["narrow_nosi"] = lambda: narrow_tiles(stress_imaging=False)
# 1000 columns: the quarter tiles [791, 919) are followed by a second quarter
# column masked to its first 78 columns, [919, 997) up to MX-3
GPU_EXTRA["narrow_masked"] = lambda: narrow_tiles(mx=1000)
# a PML too wide for the velocity launch's edge blocks (TPML > 124): rows 1 and
# PMLZH in a launch of their own
GPU_EXTRA["wide_pml"] = lambda: tiny_layered(steps=60, tpml=130)
