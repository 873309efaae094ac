"""GPU parity of the particle coupling against the FP64 oracle.

Bars: solid/fluid flags, owners, link sets (sorted by idx, dir; q and r bitwise) and new-fluid
masks bit-exact; particle state and fields within 1e-12 relative (bitwise expected, since both
sides sum link contributions in the same fixed order).
"""
import numpy as np
import pytest

from cases import particle_case

pytestmark = pytest.mark.gpu

P_ARRAYS = ["pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "omega_x", "omega_y",
            "omega_z", "theta_x", "theta_y", "theta_z", "force_x", "force_y", "force_z",
            "torque_x", "torque_y", "torque_z", "force_prev_x", "force_prev_y", "force_prev_z",
            "force_prev2_x", "force_prev2_y", "force_prev2_z", "torque_prev_x", "torque_prev_y",
            "torque_prev_z", "torque_prev2_x", "torque_prev2_y", "torque_prev2_z", "temp",
            "heat", "heat_prev"]


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0 and b.size == 0:
        return 0.0
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def sorted_links(L):
    idx = (L["cell_k"].astype(np.int64) * 100000 + L["cell_j"]) * 100000 + L["cell_i"]
    order = np.lexsort((L["dir"], idx))
    return L[order]


def make_pair(rb, O, kw, ppos, ptemp=None, seed=3, amp=1e-2):
    case = rb.default_case(**kw)
    sim, ora = rb.Sim(case), O.Oracle(case)
    T = rb.linear_temperature(case, seed, amp)
    ppos = np.asarray(ppos, dtype=np.float64).T.reshape(-1)  # (P,3) -> x[], y[], z[]
    sim.init(temperature=T, ppos=ppos, ptemp=ptemp)
    ora.init(temperature=T, ppos=ppos, ptemp=ptemp)
    return sim, ora


def check_geometry(sim, ora, where):
    ib_g, ow_g = sim.flags()
    ib_c, ow_c = ora.flags()
    assert np.array_equal(ib_g, ib_c), f"{where}: ibnode differs"
    assert np.array_equal(ow_g, ow_c), f"{where}: owner differs"
    lg, lc = sim.links(), sorted_links(ora.links())
    assert len(lg) == len(lc), f"{where}: {len(lg)} links vs {len(lc)}"
    for k in ("cell_i", "cell_j", "cell_k", "dir", "pid"):
        assert np.array_equal(lg[k], lc[k]), f"{where}: link {k} differs"
    for k in ("q", "rx", "ry", "rz"):
        assert np.array_equal(lg[k].view(np.uint64), lc[k].view(np.uint64)), f"{where}: link {k}"


def check_particles(sim, ora, where, tol=1e-12):
    for name in P_ARRAYS:
        a, b = sim.particles(name), ora.particles(name)
        assert rel_err(a, b) <= tol, f"{where}: particle {name} rel err {rel_err(a, b):.3e}"


def check_fields(sim, ora, where, tol=1e-12):
    for name in ("f", "g", "f_collide", "g_collide"):
        e = rel_err(sim.populations(name), ora.populations(name))
        assert e <= tol, f"{where}: {name} rel err {e:.3e}"


TWO = [(8.3, 9.1, 12.4), (15.6, 14.2, 11.7)]


def test_initial_geometry(rb, O):
    sim, ora = make_pair(rb, O, particle_case(24, 24, 24, npart=2, diam=6.0, ra=1e4, beta=0.05), TWO)
    check_geometry(sim, ora, "init")


def test_overlapping_spheres_lowest_pid_owns(rb, O):
    """Two overlapping spheres: shared nodes belong to the lower pid (links.cu:45-47)."""
    pos = [(12.0, 12.2, 11.9), (14.5, 12.6, 12.3), (12.7, 14.9, 12.1)]
    sim, ora = make_pair(rb, O, particle_case(24, 24, 24, npart=3, diam=6.0, ra=1e4, beta=0.05), pos)
    check_geometry(sim, ora, "init")
    ib, ow = sim.flags()
    assert (ow == 0).sum() > (ow == 1).sum()


@pytest.mark.parametrize("steps", [1, 3, 20])
def test_two_spheres_step_parity(rb, O, steps):
    sim, ora = make_pair(rb, O, particle_case(24, 24, 24, npart=2, diam=6.0, ra=1e4, beta=0.05), TWO)
    sim.step(steps)
    ora.step(steps)
    where = f"after {steps} steps"
    check_particles(sim, ora, where)
    check_geometry(sim, ora, where)
    nf_g, pid_g = sim.new_fluid()
    nf_c, pid_c = ora.new_fluid()
    assert np.array_equal(nf_g, nf_c) and np.array_equal(pid_g, pid_c), f"{where}: new fluid"
    check_fields(sim, ora, where)
    for name in ("rho", "ux", "uy", "uz", "phi", "fx"):
        assert rel_err(sim.field(name), ora.field(name)) <= 1e-12, f"{where}: {name}"


def test_contacts_and_walls(rb, O):
    """Spheres inside the repulsion range of each other and of the walls."""
    pos = [(4.5, 10.0, 10.0), (11.5, 10.3, 9.8), (20.2, 3.6, 10.0), (10.0, 20.0, 20.6)]
    sim, ora = make_pair(rb, O, particle_case(26, 24, 24, npart=4, diam=6.0, ra=1e4, beta=0.05), pos)
    sim.step(10)
    ora.step(10)
    check_particles(sim, ora, "contacts")
    check_geometry(sim, ora, "contacts")
    check_fields(sim, ora, "contacts")


def test_moving_spheres_many_steps(rb, O):
    """Fast heavy spheres cross many nodes: refill and link rebuild every few steps."""
    kw = particle_case(32, 24, 24, npart=3, diam=6.0, ra=1e5, beta=0.5)
    pos = [(8.0, 8.0, 8.0), (8.5, 16.0, 15.0), (20.0, 12.0, 12.0)]
    sim, ora = make_pair(rb, O, kw, pos, ptemp=np.array([0.3, -0.2, 0.0]))
    total_nf = 0
    for k in range(6):
        sim.step(10)
        ora.step(10)
        nf_g, _ = sim.new_fluid()
        nf_c, _ = ora.new_fluid()
        assert np.array_equal(nf_g, nf_c)
        total_nf += len(nf_c)
        check_particles(sim, ora, f"block {k}")
        check_geometry(sim, ora, f"block {k}")
    check_fields(sim, ora, "end")
    assert ora.unhandled_links(False) == 0 and ora.unhandled_links(True) == 0


RAGGED_POS = [(6.5, 7.0, 8.0), (30.2, 24.5, 18.6), (18.0, 14.5, 3.6), (10.0, 22.0, 12.5)]


def test_ragged_lattice_moving_spheres(rb, O):
    """Four fast, heated spheres on a 37 x 29 x 23 lattice: rows that are not whole waves, a last
    workgroup of 103 nodes whose last wave holds 39, and a sphere whose link box reaches the last
    plane and row, so that link, refill and special nodes sit in that partial wave."""
    kw = particle_case(37, 29, 23, npart=4, diam=6.0, ra=1e5, beta=0.5)
    sim, ora = make_pair(rb, O, kw, RAGGED_POS, ptemp=np.array([0.3, -0.2, 0.0, 0.1]))
    total_nf = 0
    for k in range(3):
        sim.step(15)
        ora.step(15)
        nf_g, _ = sim.new_fluid()
        nf_c, _ = ora.new_fluid()
        assert np.array_equal(nf_g, nf_c), f"block {k}: new fluid"
        total_nf += len(nf_c)
        check_particles(sim, ora, f"block {k}")
        check_geometry(sim, ora, f"block {k}")
    check_fields(sim, ora, "end")
    assert total_nf > 0
    # the link boxes reach the lattice's last node (idx n - 1 lies in the top-back row)
    ib, ow = sim.flags()
    assert int(np.max(np.nonzero(ow.reshape(23, 29, 37) >= 0)[0])) >= 20


def test_moving_spheres_row_pull(rb, O):
    """The moving-sphere case on a lattice 64 nodes wide (the FP64 step's row pull), spheres near
    the x walls so that special nodes sit at the rows' first and last nodes."""
    kw = particle_case(64, 16, 16, npart=3, diam=6.0, ra=1e5, beta=0.5)
    pos = [(3.4, 8.0, 8.0), (60.7, 8.5, 7.0), (32.0, 8.0, 4.0)]
    sim, ora = make_pair(rb, O, kw, pos, ptemp=np.array([0.3, -0.2, 0.0]))
    # the first steps refill nodes of the first and last column (ix = 0, 63): the populations
    # those nodes pull through the thermal x wall come from the special list itself
    wall_refills = 0
    for s in range(10):
        sim.step(1)
        ora.step(1)
        nf_g, _ = sim.new_fluid()
        nf_c, _ = ora.new_fluid()
        assert np.array_equal(nf_g, nf_c), f"step {s + 1}: new fluid"
        wall_refills += int(np.count_nonzero(np.isin(nf_c % 64, (0, 63))))
        if wall_refills:
            check_fields(sim, ora, f"step {s + 1}")
    assert wall_refills > 0
    for k in range(1, 3):
        sim.step(10)
        ora.step(10)
        check_particles(sim, ora, f"block {k}")
        check_geometry(sim, ora, f"block {k}")
    check_fields(sim, ora, "end")


def test_reference_placement(rb, O):
    """glibc srand/rand rejection placement (pinit.cu:97-144) gives the oracle's positions."""
    kw = particle_case(40, 40, 40, npart=12, diam=6.0, ra=1e4, beta=0.05)
    case = rb.default_case(**kw)
    sim, ora = rb.Sim(case), O.Oracle(case)
    sim.init()
    ora.init()
    for name in ("pos_x", "pos_y", "pos_z", "temp"):
        assert np.array_equal(sim.particles(name), ora.particles(name))
    check_geometry(sim, ora, "placed")


def _pairs_layout(rng, nx, ny, nz, spacing, split, origin=5.0):
    pos = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = np.array([origin + i * spacing, origin + j * spacing, origin + k * spacing])
                c = c + rng.uniform(-0.3, 0.3, 3)
                d = rng.normal(size=3)
                d *= split / (2.0 * np.linalg.norm(d))
                pos.append(tuple(c + d))
                pos.append(tuple(c - d))
    return pos


def test_many_spheres_cell_grid(rb, O):
    """48 spheres in close pairs: the pair search runs through the cell grid."""
    rng = np.random.default_rng(7)
    pos = _pairs_layout(rng, 4, 3, 2, 8.0, 3.6)
    kw = particle_case(40, 32, 24, npart=len(pos), diam=3.0, ra=1e4, beta=0.05)
    sim, ora = make_pair(rb, O, kw, pos)
    sim.step(4)
    ora.step(4)
    check_particles(sim, ora, "cell grid")
    check_geometry(sim, ora, "cell grid")
    check_fields(sim, ora, "cell grid")
    assert np.count_nonzero(ora.particles("force_x")) == len(pos)


def test_crowded_cell_falls_back_to_all_pairs(rb, O):
    """A cluster of 20 overlapping small spheres overflows one cell of the grid."""
    rng = np.random.default_rng(11)
    # one grid cell (width >= 1.01 * (2R + REP_MIN_GAP) = 3.23) holds all 20 centres
    cluster = [tuple(np.array([11.5, 8.9, 8.6]) + rng.uniform(-0.5, 0.5, 3)) for _ in range(20)]
    spread = [tuple(rng.uniform([4, 4, 4], [36, 28, 20])) for _ in range(20)]
    pos = cluster + spread
    kw = particle_case(40, 32, 24, npart=len(pos), diam=1.2, ra=1e4, beta=0.05)
    sim, ora = make_pair(rb, O, kw, pos)
    sim.step(2)
    ora.step(2)
    check_particles(sim, ora, "crowded")
    check_geometry(sim, ora, "crowded")
    check_fields(sim, ora, "crowded")


@pytest.mark.parametrize("side,spacing", [(4, 2.4), (5, 1.9)])
def test_dense_cluster_candidates(rb, O, side, spacing):
    """A side^3 block of small spheres, no cell over its capacity: the spheres near the block's
    centre see 64 (side 4) or 125 (side 5) pair candidates in their 27 cells, the first within
    the candidate list, the second beyond it (the all-pairs fallback); both agree with the
    oracle bit for bit."""
    centre = np.array([20.0, 16.0, 12.0])
    offs = (np.arange(side) - (side - 1) / 2.0) * spacing
    rng = np.random.default_rng(side)
    pos = [tuple(centre + np.array([a, b, c]) + rng.uniform(-0.05, 0.05, 3))
           for a in offs for b in offs for c in offs]
    kw = particle_case(40, 32, 24, npart=len(pos), diam=1.2, ra=1e4, beta=0.05)
    sim, ora = make_pair(rb, O, kw, pos)
    sim.step(2)
    ora.step(2)
    check_particles(sim, ora, f"cluster {side}")
    check_geometry(sim, ora, f"cluster {side}")
    check_fields(sim, ora, f"cluster {side}")


def test_particles_enabled_with_none(rb, O):
    """ACTIVATE_PARTICLES with a particle count of zero runs the single-phase step."""
    kw = particle_case(20, 16, 12, npart=0, diam=4.0, ra=1e5, beta=0.05)
    case = rb.default_case(**kw)
    assert rb.derive(case).npart == 0
    sim, ora = rb.Sim(case), O.Oracle(case)
    T = rb.linear_temperature(case, 4, 1e-2)
    sim.init(temperature=T)
    ora.init(temperature=T)
    sim.step(6)
    ora.step(6)
    check_fields(sim, ora, "no particles")
    assert len(sim.links()) == 0


# Bounds about 10x above the deviations from FP64 measured after 300 steps of the moving-sphere
# case below (FP32: position 3.7e-6, velocity 9.4e-7, temperature 2.2e-7, force 1.4e-6, phi
# 1.0e-6; FP16 storage: 3.6e-5, 1.8e-5, 8.3e-6, 5.8e-4, 1.1e-4; BF16 storage: 7.9e-4, 2.3e-4,
# 1.2e-4, 3.0e-2, 7.1e-3).  BF16's figures moved by up to 10x between two builds whose FP32
# arithmetic differs only in the compiler's fusing of products (force 1.7e-3 -> 3.0e-2, phi
# 8.3e-4 -> 7.1e-3), so its bounds sit 10x above the larger.  Position in lattice units, the
# rest relative to the FP64 maximum.
LOW_PRECISION_PARTICLE_BOUNDS = {
    1: dict(pos=4e-5, vel=1e-5, temp=3e-6, force=2e-5, phi=1e-5),
    2: dict(pos=4e-4, vel=2e-4, temp=1e-4, force=6e-3, phi=1.2e-3),
    3: dict(pos=8e-3, vel=2.5e-3, temp=1.2e-3, force=3e-1, phi=7e-2),
}


@pytest.mark.parametrize("prec", [1, 2, 3])
def test_low_precision_particles_track_fp64(rb, prec):
    """FP32 arithmetic with FP32, FP16 or BF16 storage against the FP64 form on four moving,
    heated spheres: the particle phase (IBB, sums, motion, refill) runs in FP64 in every form,
    so the deviations come from the stored populations alone and stay within the bounds above
    over 300 steps."""
    kw = particle_case(40, 32, 32, npart=4, diam=6.0, ra=1e5, beta=0.5)
    pos = [(10.0, 10.0, 10.0), (28.5, 16.0, 15.0), (20.0, 22.0, 22.0), (12.0, 24.0, 9.0)]
    ppos = np.asarray(pos, dtype=np.float64).T.reshape(-1)
    ptemp = np.array([0.3, -0.2, 0.0, 0.1])
    sims = {}
    for p in (0, prec):
        case = rb.default_case(**kw, precision=p)
        sims[p] = rb.Sim(case)
        sims[p].init(temperature=rb.linear_temperature(case, 3, 1e-2), ppos=ppos, ptemp=ptemp)
        sims[p].step(300)
    ref, low = sims[0], sims[prec]
    bound = LOW_PRECISION_PARTICLE_BOUNDS[prec]
    dpos = max(float(np.max(np.abs(low.particles(k) - ref.particles(k))))
               for k in ("pos_x", "pos_y", "pos_z"))
    assert dpos <= bound["pos"], dpos
    vref = [ref.particles(k) for k in ("vel_x", "vel_y", "vel_z")]
    vmax = max(float(np.max(np.abs(v))) for v in vref)
    dvel = max(float(np.max(np.abs(low.particles(k) - v)))
               for k, v in zip(("vel_x", "vel_y", "vel_z"), vref)) / vmax
    assert dvel <= bound["vel"], dvel
    assert rel_err(low.particles("temp"), ref.particles("temp")) <= bound["temp"]
    assert rel_err(low.particles("force_x"), ref.particles("force_x")) <= bound["force"]
    assert rel_err(low.field("phi"), ref.field("phi")) <= bound["phi"]
