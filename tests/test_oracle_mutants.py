"""The oracle's pins catch the mistakes they are meant to catch.

oracle.cpp carries numbered planted mistakes (ORACLE_MUTANT = k, none in a normal build).  Each
test here builds the oracle with one of them and runs the pin that targets it (in
tests/test_oracle_pins.py or tests/test_oracle.py) against that build: the pin must fail.  The
unmutated oracle built the same way must pass the same pins (the control), so a failure comes
from the mistake and not from the build.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "oracle.cpp")
PINS = os.path.join(ROOT, "tests", "test_oracle_pins.py")
SELF = os.path.join(ROOT, "tests", "test_oracle.py")

IBB_F = f"{PINS}::test_ibb_f_bouzidi_linear_fields"
IBB_G = f"{PINS}::test_ibb_g_bouzidi_linear_fields_with_density"
NEW_FLUID = f"{PINS}::test_new_fluid_predicate_brute_force"
NU = f"{PINS}::test_nu_with_particle_nodes_closed_form"
KED = f"{PINS}::test_ked_ted_closed_form"
OVERWRITE = f"{PINS}::test_output_overwrite_means_profile_closed_form"
MACROVAR = f"{SELF}::test_macrovar_moments"
COLLIDE_F = f"{SELF}::test_collision_f_conserves_mass_and_adds_force"
COLLIDE_G = f"{SELF}::test_collision_g_moments"
STREAM_F = f"{SELF}::test_streaming_delta_and_walls"
STREAM_G = f"{SELF}::test_thermal_wall_rules"
LINKS_Q = f"{SELF}::test_links_axis_aligned_q"
LINKS_BF = f"{SELF}::test_links_brute_force"
FORCE = f"{SELF}::test_moving_sphere_first_step_force_closed_form"
HEAT = f"{SELF}::test_sphere_heat_closed_form"
REPULSION = f"{SELF}::test_repulsion_closed_form"
MOVE = f"{SELF}::test_move_closed_form"
REFILL_FALLBACK = f"{SELF}::test_refill_fallback_moments_closed_form"
REFILL_LINEAR = f"{SELF}::test_refill_extrapolation_exact_for_linear_fields"
REFILL_MOMENTS = f"{SELF}::test_refill_writes_consistent_moments"
INIT = f"{SELF}::test_init_uniform_temperature"

# mutant number (oracle.cpp) -> (the mistake, the pin that must catch it)
MUTANTS = {
    1: ("bounce_back_particles: x and x - c swapped (q < 1/2)", IBB_F),
    2: ("bounce_back_particles: i and opp swapped (q < 1/2)", IBB_F),
    3: ("bounce_back_particles: i and opp swapped (q >= 1/2)", IBB_F),
    4: ("bounce_back_particles: momentum term dropped", IBB_F),
    5: ("bounce_back_thermal: x and x - c swapped (q <= 1/2)", IBB_G),
    6: ("bounce_back_thermal: i and opp swapped (q > 1/2)", IBB_G),
    7: ("bounce_back_thermal: (1 + rho) factor dropped", IBB_G),
    8: ("bounce_back_thermal: ibnode in place of ibnode_prev", IBB_G),
    9: ("new fluid: < for <=", NEW_FLUID),
    10: ("new fluid: current owner in place of the previous one", NEW_FLUID),
    11: ("new fluid: previous and current flags swapped", NEW_FLUID),
    12: ("output_nu: previous flags", NU),
    13: ("output_nu: wall sums normalised by the fluid count", NU),
    14: ("output_nu: actual solid fraction in place of PARTICLE_FRACTION", NU),
    15: ("compute_Ked: force term dropped", KED),
    16: ("compute_ted: denominator with rho + 1", KED),
    17: ("overwrite_particle_nodes: rotation dropped", OVERWRITE),
    18: ("macrovar: u without F_old/2", MACROVAR),
    19: ("collision_BGK: Guo term (c_x - u_x) reversed", COLLIDE_F),
    20: ("collision_BGK: Guo coefficient 1 - 1/(2 tau) left out", COLLIDE_F),
    21: ("collision_BGK_scalar: lambda0 term of geqnew_0 with + for -", COLLIDE_G),
    22: ("collision_BGK_scalar: grad_denom with rho + 1", COLLIDE_G),
    23: ("collision_BGK_scalar: 3 phi c.F source term dropped", COLLIDE_G),
    24: ("streaming: a wall returns i instead of opp(i)", STREAM_F),
    25: ("streaming_scalar: x walls without (1 + rho)", STREAM_G),
    26: ("streaming_scalar: hot and cold walls swapped", STREAM_G),
    27: ("streaming_scalar: x walls before the z and y walls", STREAM_G),
    28: ("compute_link_q: larger root", LINKS_Q),
    29: ("marking: the highest pid keeps a shared node", LINKS_BF),
    30: ("momentum exchange with (c_opp + u_w)", FORCE),
    31: ("heat with g*_i + g_opp", HEAT),
    32: ("repulsion: gap measured from one radius", REPULSION),
    33: ("update_particles: force history weights 1, 1, 0", MOVE),
    34: ("update_particles: buoyancy with the wrong sign", MOVE),
    35: ("refill fallback: e.u divided by RT", REFILL_FALLBACK),
    36: ("refill: first-order samples only", REFILL_LINEAR),
    37: ("refill: F kept at the refilled node", REFILL_MOMENTS),
    38: ("init_g: lambda0 term of g_0 with + for -", INIT),
}
PINNED = sorted({t for _, t in MUTANTS.values()})


def _build(k, out_dir):
    """Builds oracle.cpp with ORACLE_MUTANT = k (0: none) into out_dir; returns the path."""
    so = os.path.join(out_dir, f"liboracle_mut{k}.so")
    cmd = ["g++", "-O1", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-std=c++17",
           "-fPIC", "-shared", f"-DORACLE_MUTANT={k}", "-o", so, ORACLE]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build of mutant {k} failed:\n{r.stderr}")
    return so


def _run_pins(lib, tests):
    """Runs the named pins against the oracle build lib; returns (returncode, output)."""
    env = dict(os.environ, RBTP_ORACLE_LIB=lib, OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", *tests],
                       capture_output=True, text=True, env=env, cwd=ROOT)
    return r.returncode, r.stdout + r.stderr


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    out = str(tmp_path_factory.mktemp("oracle_mutants"))
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        libs = dict(zip([0, *MUTANTS], ex.map(lambda k: _build(k, out), [0, *MUTANTS])))
        jobs = {0: ex.submit(_run_pins, libs[0], PINNED)}
        for k, (_, test) in MUTANTS.items():
            jobs[k] = ex.submit(_run_pins, libs[k], [test])
        return {k: j.result() for k, j in jobs.items()}


def test_unmutated_build_passes_the_pins(results):
    rc, out = results[0]
    assert rc == 0, out[-3000:]


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_pin_catches_planted_mistake(results, k):
    what, test = MUTANTS[k]
    rc, out = results[k]
    assert rc == 1, f"mutant {k} ({what}) was not caught by {test}:\n{out[-3000:]}"
    assert "1 failed" in out, out[-3000:]
