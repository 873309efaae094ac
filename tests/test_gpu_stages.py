"""The per-stage entry point (rbtp_stage) against the oracle's stages, one stage at a time.

Both sides start from the same state, reached by whole fused steps (GPU) and whole oracle steps,
then run the fifteen stages of main.cu in order; after every stage all of the reference's arrays
are compared: d_f, d_g, d_f_collide, d_g_collide, rho/u/phi/F_x, ibnode/owner, the link list,
the new-fluid list and the 33 particle arrays.  The bar is bit-exactness for the flags, links and
new fluid and 1e-12 relative for the values; the FP64 stages evaluate the reference's operation
order, so equality is expected and asserted for the lattice arrays.  After the staged step both
sides continue with whole steps and agree again, so the staged state hands back to the fused
kernel correctly.  A divergence is therefore reported with the stage that introduced it.
"""
import numpy as np
import pytest

from cases import particle_case, single_phase
from test_gpu_particles import P_ARRAYS, sorted_links

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0 and b.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def compare_all(sim, ora, where, particles):
    for name in ("f", "g", "f_collide", "g_collide"):
        a, b = sim.populations(name), ora.populations(name)
        assert np.array_equal(a, b), f"{where}: {name} (rel err {rel_err(a, b):.3e})"
    for name in ("rho", "ux", "uy", "uz", "phi", "fx"):
        a, b = sim.field(name), ora.field(name)
        assert np.array_equal(a, b), f"{where}: field {name} (rel err {rel_err(a, b):.3e})"
    if not particles:
        return
    ib_g, ow_g = sim.flags()
    ib_c, ow_c = ora.flags()
    assert np.array_equal(ib_g, ib_c) and np.array_equal(ow_g, ow_c), f"{where}: flags"
    lg, lc = sim.links(), sorted_links(ora.links())
    assert len(lg) == len(lc), f"{where}: {len(lg)} links vs {len(lc)}"
    for k in ("cell_i", "cell_j", "cell_k", "dir", "pid"):
        assert np.array_equal(lg[k], lc[k]), f"{where}: link {k}"
    for k in ("q", "rx", "ry", "rz"):
        assert np.array_equal(lg[k].view(np.uint64), lc[k].view(np.uint64)), f"{where}: link {k}"
    nf_g, pid_g = sim.new_fluid()
    nf_c, pid_c = ora.new_fluid()
    assert np.array_equal(nf_g, nf_c) and np.array_equal(pid_g, pid_c), f"{where}: new fluid"
    for name in P_ARRAYS:
        e = rel_err(sim.particles(name), ora.particles(name))
        assert e <= 1e-12, f"{where}: particle {name} rel err {e:.3e}"


def run_staged(rb, O, kw, ppos=None, ptemp=None, warm=3, staged_steps=2):
    case = rb.default_case(**kw)
    sim, ora = rb.Sim(case), O.Oracle(case)
    T = rb.linear_temperature(case, 7, 1e-2)
    sim.init(temperature=T, ppos=ppos, ptemp=ptemp)
    ora.init(temperature=T, ppos=ppos, ptemp=ptemp)
    particles = bool(kw.get("activate_particles")) and sim.np > 0
    sim.step(warm)
    ora.step(warm)
    refilled = 0
    for k in range(staged_steps):
        assert sim.next_stage == "macrovar"
        for st in rb.STAGES:
            sim.stage(st)
            ora.stage(st)
            compare_all(sim, ora, f"staged step {k}, after {st}", particles)
        if particles:
            refilled += len(ora.new_fluid()[0])
    sim.refilled_in_staged_steps = refilled
    assert sim.current_step == warm + staged_steps
    sim.step(4)
    ora.step(4)
    compare_all(sim, ora, "fused steps after the staged ones", particles)
    return sim, ora


def test_stages_single_phase(rb, O):
    run_staged(rb, O, single_phase(20, 16, 12, ra=1e6))


def test_stages_moving_spheres(rb, O):
    """Three spheres sinking from rest (links, a sphere near the z = 0 wall, refill): every
    particle stage is exercised, and the third staged step refills nodes."""
    kw = particle_case(32, 24, 24, npart=3, diam=6.0, ra=1e5, beta=0.5)
    pos = np.array([[8.0, 8.0, 8.0], [8.5, 16.0, 15.0], [20.0, 12.0, 4.2]]).T.reshape(-1)
    sim, ora = run_staged(rb, O, kw, ppos=pos, ptemp=np.array([0.3, -0.2, 0.0]), warm=12,
                          staged_steps=3)
    assert sim.refilled_in_staged_steps > 0  # step 15 refills 8 nodes
    assert ora.unhandled_links(False) == 0


def test_stage_order_and_form_are_enforced(rb):
    kw = single_phase(12, 10, 8)
    sim = rb.Sim(rb.default_case(**kw))
    sim.init()
    with pytest.raises(rb.RbtpError) as e:
        sim.stage("forcing")  # a staged step starts with macrovar
    assert e.value.code == -5
    sim.stage("macrovar")
    with pytest.raises(rb.RbtpError) as e:
        sim.stage("collide_f")  # forcing comes first
    assert e.value.code == -5
    with pytest.raises(rb.RbtpError) as e:
        sim.step(1)  # the staged step is not finished
    assert e.value.code == -5
    for st in rb.STAGES[1:]:
        sim.stage(st)
    sim.step(1)
    assert sim.current_step == 2
    low = rb.Sim(rb.default_case(**kw, precision=rb.PREC_FP16S))
    low.init()
    with pytest.raises(rb.RbtpError) as e:
        low.stage("macrovar")
    assert e.value.code == -5
