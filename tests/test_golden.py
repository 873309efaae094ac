"""The oracle still reproduces the committed golden fixtures bit for bit (tests/golden/: a
fluid-only lattice, one sphere, two moving spheres with refill), so that any change of the
oracle's arithmetic is seen."""
import pytest

from golden_check import FIXTURES, compare, load, replay


def test_fixtures_present():
    assert set(FIXTURES) >= {"fluid_8x8x4", "sphere_12x10x10", "moving_spheres_16x12x12"}


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_reproduces_fixture(O, name):
    fx = load(name)
    ora = replay(O, fx)
    try:
        compare(ora, fx)
    finally:
        ora.close()
