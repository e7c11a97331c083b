"""bq_trace_misfit_noalloc (no allocation, fixed device staging array) against
bq_trace_misfit, through the C ABI (needs a GPU): two calls in a row, the
second with fewer receivers and steps than the first, so that a value left in
the staging array by the first call would show."""
import numpy as np
import pytest

import jobs
from biotfd import lib

pytestmark = pytest.mark.gpu


def _pair(make, steps, **kw):
    a = lib.Simulation(make(steps=steps))
    b = lib.Simulation(make(steps=steps, f0=kw.get("f0b", 70.0)))
    a.run(steps)
    b.run(steps)
    return a, b


def test_noalloc_misfit_equals_allocating_twice_then_smaller():
    big = _pair(jobs.medium_layered, 80, f0b=55.0)     # 40 receivers
    small = _pair(jobs.tiny_layered, 30, f0b=90.0)     # 10 receivers
    assert big[0].mr > small[0].mr
    for (a, b), first, n in ((big, 1, 80), (small, 5, 20)):
        want_tot, want_per = lib.trace_misfit(a, b, first, n, per_receiver=True)
        got_tot, got_per = lib.trace_misfit(a, b, first, n, per_receiver=True,
                                            entry="bq_trace_misfit_noalloc")
        np.testing.assert_array_equal(got_tot, want_tot)
        np.testing.assert_array_equal(got_per, want_per)
        assert want_tot[0] > 0 and want_tot[1] > 0
    # totals without the per-receiver output
    a, b = small
    np.testing.assert_array_equal(lib.trace_misfit(a, b, 5, 20, entry="bq_trace_misfit_noalloc"),
                                  lib.trace_misfit(a, b, 5, 20))
