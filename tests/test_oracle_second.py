"""The oracle's second-order build option (the reference's -DSECOND: A = 0,
B = 1, mod_auxil.F90:15-21) pinned by what the mathematics fixes.

The reference's SECOND build does not compile (inh_qx_int_0.F90:47 and
inh_v_uv_int_0.F90:48 leave a parenthesis unclosed), so there is no reference
fixture. One FP64 oracle step on a homogeneous grid with TXX a cubic in x
shows which x-difference the interior velocity update takes: the 4th-order
pair (A, B) = (-1/24, 9/8) is exact for cubics at the half-cell point, the
2-point difference of the SECOND build is not, and both are checked against
their closed forms. On the explicit surface's row 0 the SECOND form adds
TXZ(i,1) alone.
"""
import numpy as np
import pytest

import jobs
import oracle
from biotfd import hostsetup
from biotfd.formats import MO_INDEX


def run(variant, si, **fields):
    job = jobs.tiny_homog(steps=4)
    job.stress_imaging = si
    job.variant = variant
    job.z_source = 6.0 * job.h
    job.stf_v = np.zeros_like(job.stf_v)
    job.receivers = []
    sim = oracle.OracleSim(job, "f64")
    nz, nx = hostsetup.grid_shape(job)
    x = np.arange(nx) + (1 - job.tpml)          # x index i of each column
    z = np.arange(nz)
    for name, f in fields.items():
        sim.set_field(name, f(x[None, :], z[:, None]) * np.ones((nz, nx)))
    sim.run(1)
    return job, sim, hostsetup.scaled_table(job, np.float64)[0]


def cubic(x):
    return 0.02 * x ** 3 - 0.3 * x ** 2 + 0.5 * x


@pytest.mark.parametrize("variant", [0, 4])
def test_interior_x_difference(variant):
    job, sim, tab = run(variant, True, TXX=lambda x, z: cubic(x))
    i = np.arange(4, job.mx - 1)                 # interior QX/U columns
    c = i + job.tpml - 1
    got = -sim.field("QX")[10, c] / tab[MO_INDEX["REL1_QX"]]
    if variant == 4:
        want = cubic(i) - cubic(i - 1)           # 2-point difference
    else:
        want = 0.06 * (i - 0.5) ** 2 - 0.6 * (i - 0.5) + 0.5   # exact derivative at i - 1/2
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_surface_row0_second_form():
    """Row 0 of the explicit surface, SECOND: QX = -REL1_QX ((TXX(i)-TXX(i-1)) + TXZ(i,1))."""
    job, sim, tab = run(4, False, TXX=lambda x, z: cubic(x), TXZ=lambda x, z: 0.7 * z + 0.1 * z ** 2)
    i = np.arange(4, job.mx - 1)
    c = i + job.tpml - 1
    got = -sim.field("QX")[0, c] / tab[MO_INDEX["REL1_QX"]]
    np.testing.assert_allclose(got, cubic(i) - cubic(i - 1) + 0.8, rtol=1e-12, atol=1e-12)
