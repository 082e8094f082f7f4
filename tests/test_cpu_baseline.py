"""The benchmark's all-core CPU baseline is the oracle built with OpenMP
(oracle/Makefile target mt).  It must do the same work as the single-thread
checker: the same particles, cells, order and cell sums bit for bit, and the
wall sums (added atomically in thread order) to round-off."""
import os

import numpy as np
import pytest

import helpers as H
import oracle_bind as ob

MT = os.path.join(ob.ROOT, "oracle", "_build", "libdsmc_oracle_mt.so")


@pytest.mark.skipif(not os.path.exists(MT), reason="threaded oracle not built (make -C oracle mt)")
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_threaded_oracle_does_the_same_steps(threads):
    ob.lib(MT).or_set_threads(threads)
    assert ob.lib(MT).or_threads() == threads
    case = H.small_cylinder(n=60, ppc=10, flags=H.abi.HS_FLAG_SAMPLE | H.abi.HS_FLAG_WALL)
    a = ob.OracleState(case)
    b = ob.OracleState(case, lib_path=MT)
    H.fill_cylinder(a, case)
    H.fill_cylinder(b, case)
    a.evolve(30)
    b.evolve(30)
    H.assert_same_particles(b.particles(), a.particles(), f"threads={threads}")
    np.testing.assert_array_equal(b.sample_array(), a.sample_array())
    np.testing.assert_array_equal(b.sigma, a.sigma)
    np.testing.assert_array_equal(b.rem, a.rem)
    np.testing.assert_allclose(b.wall, a.wall, rtol=1e-12, atol=1e-300)
    assert b.counters()[:7] == a.counters()[:7]
