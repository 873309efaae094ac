"""rbtp_nu_noalloc (static device scratch, no allocation) against rbtp_nu, through the C-ABI.

Two calls in a row on different lattices, the second smaller than the first, so that the second
call runs on scratch that still holds the first call's larger contents.  The two entry points
share their kernels and the combination is the same sequence of IEEE operations, so the bar is
bitwise equality (within 1e-10 of Nu is the written tolerance of the Nu parity tests).
"""

import numpy as np
import pytest

from cases import particle_case, single_phase

pytestmark = pytest.mark.gpu


def nu_pair(rb, sim):
    a = np.zeros(7)
    b = np.zeros(7)
    assert rb.lib.rbtp_nu(sim._h, a) == 0, rb.lib.rbtp_last_error()
    assert rb.lib.rbtp_nu_noalloc(sim._h, b) == 0, rb.lib.rbtp_last_error()
    return a, b


def test_noalloc_nu_matches_nu_on_two_calls(rb):
    big = rb.Sim(rb.default_case(**particle_case(48, 32, 24, npart=3, diam=6.0, ra=1e5,
                                                 beta=0.5)))
    T = rb.linear_temperature(big.case, 8, 1e-2)
    big.init(temperature=T, ppos=np.array([[12.0, 30.0, 20.0], [10.0, 16.0, 20.0],
                                           [8.0, 12.0, 14.0]]).reshape(-1))
    big.step(40)
    a, b = nu_pair(rb, big)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (a, b)
    assert np.all(np.abs(a - b) <= 1e-10 * np.maximum(1.0, np.abs(a)))
    assert a[0] == 40 and a[2] != 0.0

    small = rb.Sim(rb.default_case(**single_phase(20, 16, 12, ra=1e6)))
    small.init(temperature=rb.linear_temperature(small.case, 9, 1e-2))
    small.step(15)
    a2, b2 = nu_pair(rb, small)
    assert np.array_equal(a2.view(np.uint64), b2.view(np.uint64)), (a2, b2)
    assert not np.array_equal(a2, a)

    # and back on the larger lattice after the smaller one
    big.step(5)
    a3, b3 = nu_pair(rb, big)
    assert np.array_equal(a3.view(np.uint64), b3.view(np.uint64))
    big.close()
    small.close()


def test_noalloc_nu_rejects_lattices_beyond_its_scratch(rb):
    sim = rb.Sim(rb.default_case(**single_phase(8193, 3, 3, ra=1e5)))
    sim.init()
    sim.step(1)
    out = np.zeros(7)
    rc = rb.lib.rbtp_nu_noalloc(sim._h, out)
    assert rc == -6
    assert b"scratch limit" in rb.lib.rbtp_last_error()
    sim.close()
