"""The oracle's explicit free surface (no stress imaging), pinned by what the
mathematics fixes.

The reference cannot run this variant (its call of INH_QZ_2ND_0 faults, see
tests/golden/make_golden.py), so no reference fixture exists for it. Instead
every one-sided surface stencil of the restatement is driven through one step
of the FP64 oracle on a homogeneous grid whose only non-zero field is a
polynomial in depth, and the derivative it yields is compared with the exact
one. On the staggered grid W, QZ, TXZ of row l sit at depth l and U, QX, TZZ,
PRES, TXX at l + 1/2 (cells); TZZ, PRES and TXZ vanish at the surface z = 0.
Each stencil is exact for the degree it is built for, so a wrong coefficient,
sign or row index fails here.
"""
import numpy as np
import pytest

import jobs
import oracle
from biotfd import hostsetup
from biotfd.formats import MO_INDEX

A, B, C3 = 0.7, -0.3, 0.05          # g(z) = A z + B z^2 + C3 z^3 (+ D4 z^4), g(0) = 0
D4 = 0.011


def g(z, deg):
    z = np.asarray(z, np.float64)
    out = A * z + B * z ** 2
    if deg >= 3:
        out = out + C3 * z ** 3
    if deg >= 4:
        out = out + D4 * z ** 4
    return out


def dg(z, deg):
    out = A + 2 * B * z
    if deg >= 3:
        out = out + 3 * C3 * z ** 2
    if deg >= 4:
        out = out + 4 * D4 * z ** 3
    return out


def one_step(variant, **fields):
    """One oracle step (FP64) of a homogeneous non-imaging grid from the given
    initial fields (functions of depth in cells), with a silent source."""
    job = jobs.tiny_homog(steps=4)
    job.stress_imaging = False
    job.variant = variant
    job.z_source = 6.0 * job.h
    job.stf_v = np.zeros_like(job.stf_v)
    job.receivers = []
    sim = oracle.OracleSim(job, "f64")
    nz, nx = hostsetup.grid_shape(job)
    for name, prof in fields.items():
        sim.set_field(name, np.repeat(prof(np.arange(nz))[:, None], nx, axis=1))
    sim.run(1)
    tab = hostsetup.scaled_table(job, np.float64)[0]
    return job, sim, tab


def cols(job, lo, hi):
    """Array columns of x indices lo..hi."""
    off = job.tpml - 1
    return slice(lo + off, hi + off + 1)


@pytest.mark.parametrize("field,r1,r3", [("TZZ", "REL1_QZ", "REL1_W"), ("P", "REL2_QZ", "REL3_W")])
def test_row0_row1_qz_w_interior_cubic(field, r1, r3):
    """INH_QZ_INT_0 / INH_V_W_INT_0 (35/8, -35/24, 21/40, -5/56) and
    INH_V_QZ_INT_1 / INH_V_W_INT_1 (-31/24, 29/24, -3/40, 1/168): the
    derivative at depth 0 and 1 of a cubic that vanishes at the surface."""
    job, sim, tab = one_step(0, **{field: lambda l: g(l + 0.5, 3)})
    c = cols(job, 3, job.mx - 2)
    qz, w = sim.field("QZ"), sim.field("W")
    for row in (0, 1):
        want = dg(row, 3)
        np.testing.assert_allclose(-qz[row, c] / tab[MO_INDEX[r1]], want, rtol=1e-12)
        np.testing.assert_allclose(w[row, c] / tab[MO_INDEX[r3]], want, rtol=1e-12)


@pytest.mark.parametrize("field,r1,r3", [("TZZ", "REL1_QZ", "REL1_W"), ("P", "REL2_QZ", "REL3_W")])
def test_row0_qz_w_side_strips_quadratic(field, r1, r3):
    """INH_QZ_2ND_0 / INH_V_W_2ND_0 (3, -1/3): exact for a quadratic that
    vanishes at the surface, in both x-PML strips."""
    job, sim, tab = one_step(0, **{field: lambda l: g(l + 0.5, 2)})
    qz, w = sim.field("QZ"), sim.field("W")
    for lo, hi in ((2 - job.tpml, 2), (job.mx - 1, job.mx - 1 + job.tpml)):
        c = cols(job, lo, hi)
        np.testing.assert_allclose(-qz[0, c] / tab[MO_INDEX[r1]], dg(0, 2), rtol=1e-12)
        np.testing.assert_allclose(w[0, c] / tab[MO_INDEX[r3]], dg(0, 2), rtol=1e-12)


def test_row0_qx_u_interior_cubic_and_strips_quadratic():
    """INH_QX_INT_0 / INH_V_UV_INT_0 (17/24, 3/8, -5/24, 1/24 of TXZ rows 1-4):
    derivative at depth 1/2 of a cubic TXZ vanishing at the surface; in the
    strips (INH_QX_2ND_0 / INH_V_U_2ND_0) the 2-point difference, exact for a
    quadratic."""
    for deg in (3, 2):
        job, sim, tab = one_step(0, TXZ=lambda l: g(l, deg))
        qx, u = sim.field("QX"), sim.field("U")
        spans = [(4, job.mx - 2)] if deg == 3 else [(3 - job.tpml, 3), (job.mx - 1, job.mx - 1 + job.tpml)]
        for lo, hi in spans:
            c = cols(job, lo, hi)
            np.testing.assert_allclose(-qx[0, c] / tab[MO_INDEX["REL1_QX"]], dg(0.5, deg), rtol=1e-12)
            np.testing.assert_allclose(u[0, c] / tab[MO_INDEX["REL1_U"]], dg(0.5, deg), rtol=1e-12)


@pytest.mark.parametrize("variant,deg,lo", [(0, 2, None), (1, 4, 2)])
def test_row0_strain_zz(variant, deg, lo):
    """STRAIN_UW_ZZ_0: the 2-point difference (exact for quadratics) or, with
    AFDA_ZZ, the 5-point one-sided form (exact for quartics), at depth 1/2."""
    job, sim, tab = one_step(variant, W=lambda l: g(l, deg) + 0.25)
    lo = 2 - job.tpml if lo is None else lo
    hi = job.mx - 1 + (job.tpml if not variant else 0)
    c = cols(job, lo, hi)
    ezz = sim.field("TZZ")[0, c] / (np.float64(job.dt) * tab[MO_INDEX["SIG_ZZ_2"]])
    np.testing.assert_allclose(ezz, dg(0.5, deg), rtol=1e-12)


@pytest.mark.parametrize("variant,prof,want", [
    (0, lambda z: g(z, 2), dg(1.0, 2)),
    (2, lambda z: B * z ** 2 + C3 * z ** 3, 2 * B + 3 * C3),   # U'(0) = 0 = -W_x(0)
])
def test_row1_strain_xz(variant, prof, want):
    """STRAIN_UW_XZ_1 at depth 1 from U at depths 1/2, 3/2, ...: the 2-point
    difference, or with AFDA_XZ the 4-point form corrected by the surface
    condition U_z(0) = -W_x(0) (exact for cubics that satisfy it)."""
    job, sim, tab = one_step(variant, U=lambda l: prof(l + 0.5))
    c = cols(job, 3, job.mx - 1)
    exz = sim.field("TXZ")[1, c] / (np.float64(job.dt) * tab[MO_INDEX["MU"]])
    np.testing.assert_allclose(exz, want, rtol=1e-12)
