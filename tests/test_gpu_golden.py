"""The HIP library reproduces the committed golden fixtures (tests/golden/): populations, flags,
link records and new-fluid lists bit for bit, fields and particle state within 1e-12 relative,
Nu within 1e-10 (the device reductions use their own fixed order)."""
import pytest

from golden_check import FIXTURES, compare, load, replay

pytestmark = pytest.mark.gpu


def test_fixtures_present():
    assert set(FIXTURES) >= {"fluid_8x8x4", "sphere_12x10x10", "moving_spheres_16x12x12"}


@pytest.mark.parametrize("name", FIXTURES)
def test_library_reproduces_fixture(rb, name):
    fx = load(name)
    sim = replay(rb, fx)
    try:
        compare(sim, fx, field_tol=1e-12, nu_tol=1e-10)
    finally:
        sim.close()
