"""GPU parity at the benchmark configurations' own shapes, against the FP64 oracle.

  320x320x160, Ra=1e8, FP64 (the benchmark's headline workload): 4 steps from the benchmark's
      initial state, every population and field bitwise (1e-12 relative is the written bar);
  320x320x160 + 300 spheres of diameter 10 (reference placement): solid/fluid flags, owners and
      the 861k-link set bit-exact at the start and after 3 steps, new-fluid lists bit-exact,
      particle state and populations within 1e-12;
  512x512x256, Ra=1e9 + 5000 spheres of diameter 8: the same after 1 step (9.1 M links, 67 M
      nodes; about 63 GB of host memory for the oracle);
  320x320x160 in FP16 storage: against the FP64 form over 2000 steps, Nu within the 1 % bar.

The oracle runs on the host cores (OpenMP), so these tests take tens of seconds each.
Populations are compared one population at a time to keep the host memory bounded.
"""
import numpy as np
import pytest

from cases import config2, config3, config5, local_particle_temperature
from test_gpu_particles import P_ARRAYS, sorted_links

pytestmark = pytest.mark.gpu


def compare_populations(sim, ora, name, tol=1e-12, bitwise=False):
    a = sim.populations(name)
    b = ora.populations(name, copy=False)
    scale = 0.0
    diff = 0.0
    equal = True
    for i in range(27):
        scale = max(scale, float(np.max(np.abs(b[i]))))
        d = np.abs(a[i] - b[i])
        diff = max(diff, float(np.max(d)))
        equal = equal and bool(np.array_equal(a[i], b[i]))
    del a
    rel = diff / max(scale, 1e-300)
    assert rel <= tol, f"{name}: rel err {rel:.3e}"
    if bitwise:
        assert equal, f"{name}: not bitwise (max diff {diff:.3e})"
    return rel


def compare_geometry(sim, ora, where):
    ib_g, ow_g = sim.flags()
    ib_c, ow_c = ora.flags()
    assert np.array_equal(ib_g, ib_c), f"{where}: ibnode"
    assert np.array_equal(ow_g, ow_c), f"{where}: owner"
    del ib_g, ow_g, ib_c, ow_c
    lg, lc = sim.links(), sorted_links(ora.links())
    assert len(lg) == len(lc), f"{where}: {len(lg)} links vs {len(lc)}"
    for k in ("cell_i", "cell_j", "cell_k", "dir", "pid"):
        assert np.array_equal(lg[k], lc[k]), f"{where}: link {k}"
    for k in ("q", "rx", "ry", "rz"):
        assert np.array_equal(lg[k].view(np.uint64), lc[k].view(np.uint64)), f"{where}: link {k}"
    return len(lg)


def compare_particles(sim, ora, where, tol=1e-12):
    for name in P_ARRAYS:
        a, b = sim.particles(name), ora.particles(name)
        scale = max(float(np.max(np.abs(b))), 1e-300)
        err = float(np.max(np.abs(a - b))) / scale
        assert err <= tol, f"{where}: particle {name} rel err {err:.3e}"


def compare_new_fluid(sim, ora, where):
    nf_g, pid_g = sim.new_fluid()
    nf_c, pid_c = ora.new_fluid()
    assert np.array_equal(nf_g, nf_c) and np.array_equal(pid_g, pid_c), f"{where}: new fluid"
    return len(nf_c)


def compare_fields(sim, ora, where, bitwise=True):
    for name in ("rho", "ux", "uy", "uz", "phi", "fx"):
        a, b = sim.field(name), ora.field(name)
        scale = max(float(np.max(np.abs(b))), 1e-300)
        assert float(np.max(np.abs(a - b))) / scale <= 1e-12, f"{where}: {name}"
        if bitwise:
            assert np.array_equal(a, b), f"{where}: {name} not bitwise"


def particle_pair(rb, O, kw):
    """The case on both sides from the same start: the reference's placement (glibc rand,
    computed once on the host), spheres at the fluid temperature of their centre cell."""
    case = rb.default_case(**kw)
    ppos = rb.place_particles(case).reshape(-1)
    T = rb.linear_temperature(case, 2024, 1e-3)
    ptemp = local_particle_temperature(T, ppos, case.lx, case.ly, case.lz)
    sim = rb.Sim(case)
    sim.init(temperature=T, ppos=ppos, ptemp=ptemp)
    ora = O.Oracle(case)
    ora.init(temperature=T, ppos=ppos, ptemp=ptemp)
    return sim, ora


def test_config2_production_shape_bitwise(rb, O):
    """320x320x160, Ra=1e8, Pr=0.71, FP64: the benchmark's own workload, 4 steps."""
    case = rb.default_case(**config2(precision=0))
    T = rb.linear_temperature(case, 2024, 1e-3)
    sim = rb.Sim(case)
    sim.init(temperature=T)
    ora = O.Oracle(case)
    ora.init(temperature=T)
    sim.step(4)
    ora.step(4)
    for name in ("f", "g", "f_collide", "g_collide"):
        compare_populations(sim, ora, name, bitwise=True)
    compare_fields(sim, ora, "c2")
    nu_g, nu_c = sim.nu(), ora.nu()
    assert np.all(np.abs(nu_g - nu_c) <= 1e-10 * np.maximum(1.0, np.abs(nu_c))), (nu_g, nu_c)
    sim.close()
    ora.close()


def test_config3_300_spheres(rb, O):
    """320x320x160 + 300 spheres (d = 10, reference placement): geometry and links bit-exact
    at the start and after 3 steps, new fluid bit-exact, state within 1e-12."""
    sim, ora = particle_pair(rb, O, config3(precision=0))
    assert sim.np == ora.np == 300
    nlinks = compare_geometry(sim, ora, "c3 start")
    assert nlinks > 800_000
    sim.step(3)
    ora.step(3)
    compare_particles(sim, ora, "c3 step 3")
    compare_geometry(sim, ora, "c3 step 3")
    compare_new_fluid(sim, ora, "c3 step 3")
    for name in ("f", "g", "f_collide", "g_collide"):
        compare_populations(sim, ora, name)
    compare_fields(sim, ora, "c3 step 3", bitwise=False)
    sim.close()
    ora.close()


def test_config5_5000_spheres(rb, O):
    """512x512x256, Ra=1e9 + 5000 spheres (d = 8, reference placement): geometry and the
    9.1 M-link set bit-exact at the start and after one step, new fluid bit-exact, particle
    state and populations within 1e-12."""
    sim, ora = particle_pair(rb, O, config5(precision=0))
    assert sim.np == ora.np == 5000
    nlinks = compare_geometry(sim, ora, "c5 start")
    assert nlinks > 9_000_000
    sim.step(1)
    ora.step(1)
    compare_particles(sim, ora, "c5 step 1")
    compare_geometry(sim, ora, "c5 step 1")
    compare_new_fluid(sim, ora, "c5 step 1")
    for name in ("f", "g"):
        compare_populations(sim, ora, name)
    compare_fields(sim, ora, "c5 step 1", bitwise=False)
    sim.close()
    ora.close()


def test_config2_fp16_storage_tracks_fp64(rb):
    """320x320x160, Ra=1e8 in FP16 storage (FP32 arithmetic, shifted populations, shift profile
    refreshed every 500 steps) against the FP64 form, which the test above holds bitwise to
    the oracle, over 2000 steps from the same initial state: the hot-wall Nu within the
    configured bar of 1 %, the temperature field within 1e-3 of its range, and the kinetic
    energy within 1 %."""
    kw64 = config2(precision=0)
    case64 = rb.default_case(**kw64)
    T = rb.linear_temperature(case64, 2024, 1e-3)
    ref = rb.Sim(case64)
    ref.init(temperature=T)
    ref.step(2000)
    nu_r, ke_r, phi_r = ref.nu(), ref.kinetic_energy(), ref.field("phi")
    ref.close()
    low = rb.Sim(rb.default_case(**config2(precision=2)))
    low.init(temperature=T)
    low.step(2000)
    nu_l, ke_l, phi_l = low.nu(), low.kinetic_energy(), low.field("phi")
    low.close()
    assert abs(nu_l[4] - nu_r[4]) <= 1e-2 * abs(nu_r[4])
    assert float(np.max(np.abs(phi_l - phi_r))) <= 1e-3 * float(np.ptp(phi_r))
    assert abs(ke_l - ke_r) <= 1e-2 * abs(ke_r)
