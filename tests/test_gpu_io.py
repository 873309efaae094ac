"""GPU tests of the output side and the restart layout.

Output fields (%09d.dat order: rho, ux, uy, uz, phi, ted, Ked, with particle nodes overwritten)
and the plane profile are compared with the oracle; files are checked for their byte layout;
a restart from fg%09d.dat + %09d.dat continues a single-phase run bit for bit, because f, g and
F_x (recomputed from the stored phi, initial.cu:207) are all exact.
"""
import os
import subprocess

import numpy as np
import pytest

from cases import particle_case, single_phase

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rel_err(a, b):
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / scale)


def pair(rb, O, kw, ppos=None, steps=7):
    case = rb.default_case(**kw)
    sim, ora = rb.Sim(case), O.Oracle(case)
    T = rb.linear_temperature(case, 8, 1e-2)
    p = None if ppos is None else np.asarray(ppos, float).T.reshape(-1)
    sim.init(temperature=T, ppos=p)
    ora.init(temperature=T, ppos=p)
    sim.step(steps)
    ora.step(steps)
    return sim, ora


@pytest.mark.parametrize("particles", [False, True])
def test_output_fields_and_profile(rb, O, particles):
    if particles:
        kw = particle_case(24, 22, 20, npart=2, diam=6.0, ra=1e5, beta=0.05)
        sim, ora = pair(rb, O, kw, ppos=[(8.2, 10.1, 9.7), (15.9, 12.4, 10.3)])
    else:
        sim, ora = pair(rb, O, single_phase(20, 18, 12, ra=1e6))
    a, b = sim.output_fields(), ora.output_fields()
    for k, name in enumerate(("rho", "ux", "uy", "uz", "phi", "ted", "Ked")):
        assert rel_err(a[k], b[k]) <= 1e-12, name
    assert rel_err(sim.profile(), ora.profile()) <= 1e-12


def test_file_layouts(rb, O, tmp_path):
    kw = particle_case(20, 16, 14, npart=2, diam=5.0, ra=1e5, beta=0.05)
    sim, ora = pair(rb, O, kw, ppos=[(6.2, 8.1, 7.7), (13.9, 8.4, 6.3)])
    n = sim.n
    fg = tmp_path / "fg000000007.dat"
    sim.write_fg(str(fg))
    raw = np.fromfile(fg, dtype="<f8")
    assert raw.size == 2 * 27 * n
    assert np.array_equal(raw[:27 * n], sim.populations("f").reshape(-1))
    assert np.array_equal(raw[27 * n:], sim.populations("g").reshape(-1))
    flow = tmp_path / "000000007.dat"
    sim.write_flow(str(flow))
    raw = np.fromfile(flow, dtype="<f8")
    assert raw.size == 7 * n
    assert np.array_equal(raw, sim.output_fields().reshape(-1))
    pos = tmp_path / "pos_000000007.txt"
    sim.write_particles(str(pos))
    lines = pos.read_text().splitlines()
    assert len(lines) == 2
    for pid, line in enumerate(lines):
        cols = line.split()
        assert len(cols) == 34 and int(cols[0]) == pid
        assert all(len(c) <= 14 and "e" in c for c in cols[1:])
        assert float(cols[1]) == pytest.approx(sim.particles("pos_x")[pid], rel=1e-6)
        assert float(cols[31]) == pytest.approx(sim.particles("temp")[pid], rel=1e-6, abs=1e-12)


def test_single_phase_restart_is_bit_continuous(rb, tmp_path):
    kw = single_phase(24, 20, 16, ra=1e6)
    case = rb.default_case(**kw)
    T = rb.linear_temperature(case, 5, 1e-2)
    full = rb.Sim(case)
    full.init(temperature=T)
    full.step(30)
    part = rb.Sim(case)
    part.init(temperature=T)
    part.step(12)
    d = tmp_path / "Ra"
    d.mkdir()
    part.write_fg(str(d / "fg000000012.dat"))
    part.write_flow(str(d / "000000012.dat"))
    again = rb.Sim(rb.default_case(**kw, continuous_steps=12))
    again.restart_load(str(d), 12)
    assert again.current_step == 12
    again.step(18)
    for name in ("f", "g"):
        assert np.array_equal(again.populations(name), full.populations(name)), name
    assert np.array_equal(again.field("phi"), full.field("phi"))


@pytest.mark.parametrize("prec", [1, 2])
def test_reduced_precision_restart_continues(rb, tmp_path, prec):
    """An FP32- or FP16-storage run written out (populations decoded to FP64) and read back
    (encoded again into the shifted storage form) continues within the storage rounding.  F_x
    is recomputed from the stored phi on load (initial.cu:207) instead of carried over in FP32,
    so the two runs differ by single roundings that the storage then rounds again: measured
    1.1e-6 (FP32) and 1.1e-3 (FP16) after 18 more steps; a broken encoding would be O(1) off."""
    kw = single_phase(24, 20, 16, ra=1e6, precision=prec)
    case = rb.default_case(**kw)
    T = rb.linear_temperature(case, 5, 1e-2)
    full = rb.Sim(case)
    full.init(temperature=T)
    full.step(30)
    part = rb.Sim(case)
    part.init(temperature=T)
    part.step(12)
    d = tmp_path / "Ra"
    d.mkdir()
    part.write_fg(str(d / "fg000000012.dat"))
    part.write_flow(str(d / "000000012.dat"))
    again = rb.Sim(rb.default_case(**kw, continuous_steps=12))
    again.restart_load(str(d), 12)
    assert again.current_step == 12
    again.step(18)
    tol = {1: 1e-5, 2: 5e-3}[prec]
    for name in ("f", "g"):
        assert rel_err(again.populations(name), full.populations(name)) <= tol, name
    assert rel_err(again.field("phi"), full.field("phi")) <= tol


def test_restart_with_particles_reads_pos_file(rb, tmp_path):
    kw = particle_case(20, 16, 14, npart=2, diam=5.0, ra=1e5, beta=0.05)
    case = rb.default_case(**kw)
    sim = rb.Sim(case)
    sim.init(temperature=rb.linear_temperature(case, 1, 1e-2),
             ppos=np.array([6.2, 13.9, 8.1, 8.4, 7.7, 6.3]))
    sim.step(9)
    d = tmp_path / "dir"
    (d / "pos").mkdir(parents=True)
    sim.write_fg(str(d / "fg000000009.dat"))
    sim.write_flow(str(d / "000000009.dat"))
    sim.write_particles(str(d / "pos" / "pos_000000009.txt"))
    again = rb.Sim(case)
    again.restart_load(str(d), 9)
    txt = np.loadtxt(d / "pos" / "pos_000000009.txt")
    assert np.array_equal(again.particles("pos_x"), txt[:, 1])
    assert np.array_equal(again.particles("temp"), txt[:, 31])
    ib_a, ow_a = again.flags()
    ib_b, ow_b = sim.flags()
    assert abs(int(ib_a.sum()) - int(ib_b.sum())) <= 4  # positions kept to 7 digits
    again.step(5)
    assert np.all(np.isfinite(again.populations("f")))


def test_restart_missing_or_short_files(rb, tmp_path):
    kw = single_phase(12, 10, 8)
    sim = rb.Sim(rb.default_case(**kw))
    with pytest.raises(rb.RbtpError) as e:
        sim.restart_load(str(tmp_path), 5)
    assert e.value.code == -4
    np.zeros(27 * 960 + 5).tofile(tmp_path / "fg000000005.dat")
    with pytest.raises(rb.RbtpError) as e:
        sim.restart_load(str(tmp_path), 5)
    assert "corrupted" in str(e.value)


def test_rb3d_driver_outputs_and_restart(tmp_path):
    exe = os.path.join(ROOT, "rb-thermoparticles_amd", "rb3d")
    params = tmp_path / "parameters.h"
    body = """#define CONTINUOUS_STEPS {cs}
#define LX 24
#define LY 20
#define LZ 16
#define NEND {nend}
#define NDIAG 10
#define NFLOWOUT 20
#define NNUOUT 10
#define NOUT_P 10
#define RAYLEIGH 1.0e5
#define PRANDTL 0.71
#define T_HOT 0.5
#define T_COLD -0.5
#define T_REF 0.0
#define ACTIVATE_PARTICLES 1
#define PARTICLE_FRACTION 0.01
#define NPART_OVERRIDE 2
#define PARTICLE_DIAM_FRAC (6.0/24.0)
#define RHO_PARTICLE_RATIO 2.0
#define BETA 0.05
"""
    params.write_text(body.format(cs=0, nend=20))
    r = subprocess.run([exe, str(params), "--temperature", "linear:3:0.01"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Step 10" in r.stdout and "Simulation completed." in r.stdout
    d = tmp_path / "Ra1.0e+05Pr0.71"
    n = 24 * 20 * 16
    assert (d / "fg000000020.dat").stat().st_size == 2 * 27 * n * 8
    assert (d / "000000020.dat").stat().st_size == 7 * n * 8
    assert len((d / "statistics.dat").read_text().splitlines()) == 2
    nu = np.loadtxt(d / "Nu_1.00e+05.txt")
    assert nu.shape == (2, 7)
    prof = (d / "Profile_1.00e+05.txt").read_text().splitlines()
    assert len(prof) == 2 * (1 + 24)
    assert (d / "pos" / "pos_000000000.txt").exists() and (d / "pos" / "pos_000000020.txt").exists()
    params.write_text(body.format(cs=20, nend=30))
    r = subprocess.run([exe, str(params)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Restarted from step: 20" in r.stdout
    assert (d / "fg000000030.dat").exists()


def test_kinetic_energy_over_fluid_nodes(rb, O):
    """rbtp_kinetic_energy is the mean of 0.5|u|^2 over the nodes that are fluid in the current
    and in the previous geometry."""
    kw = particle_case(24, 22, 20, npart=2, diam=6.0, ra=1e5, beta=0.5)
    sim, ora = pair(rb, O, kw, ppos=[(8.2, 10.1, 9.7), (15.9, 12.4, 10.3)], steps=11)
    ib_prev, _ = ora.flags()  # geometry at the end of step 11: the previous one after step 12
    sim.step(1)
    ora.step(1)
    ib, _ = sim.flags()
    assert np.array_equal(ib, ora.flags()[0])
    fluid = (ib.reshape(-1) == 0) & (ib_prev.reshape(-1) == 0)
    u = [sim.field(k).reshape(-1)[fluid] for k in ("ux", "uy", "uz")]
    expect = np.mean(0.5 * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2))
    assert sim.kinetic_energy() == pytest.approx(expect, rel=1e-12)
    assert fluid.sum() < fluid.size


def test_measurement_probes(rb):
    """The copy and access-pattern probes return a positive rate and reject bad arguments."""
    copy = rb.bandwidth_probe(64 << 20, 2)
    assert 0.0 < copy < 1e5
    sten = rb.stencil_probe(64, 48, 40, 2)
    assert 0.0 < sten < 1e5
    with pytest.raises(rb.RbtpError):
        rb.stencil_probe(1, 48, 40, 2)
    with pytest.raises(rb.RbtpError):
        rb.stencil_probe(64, 48, 40, 0)
    # FP64 pull, AA even, AA odd, plain pull; 16-bit plain pull, row pull, two nodes per thread
    for scheme in range(7):
        assert 0.0 < rb.streaming_probe(scheme, 64, 48, 40, 2) < 1e5, scheme
    with pytest.raises(rb.RbtpError):
        rb.streaming_probe(7, 64, 48, 40, 2)
    with pytest.raises(rb.RbtpError):
        rb.streaming_probe(5, 48, 48, 40, 2)  # rows of whole waves only
    with pytest.raises(rb.RbtpError):
        rb.streaming_probe(6, 63, 48, 40, 2)  # node pairs within rows only


def test_kernel_timing_over_the_timed_steps(rb):
    """rbtp_set_kernel_timing: one pair of events per lattice launch, read at the run's
    synchronisation.  Over more steps than the library's pool of event pairs (1024) every launch
    is counted, the average launch lies within the step time, and the timed run ends in the same
    state, bit for bit, as an untimed one."""
    kw = single_phase(32, 24, 16, ra=1e5)
    case = rb.default_case(**kw)
    T = rb.linear_temperature(case, 11, 1e-2)
    sims = [rb.Sim(case), rb.Sim(case)]
    for s in sims:
        s.init(temperature=T)
    sims[0].set_kernel_timing(True)
    for s in sims:
        s.step(1100, macro_last=True)
    tm = sims[0].timing()
    assert tm.launches == 1100
    assert 0.0 < tm.kernel_ms <= tm.step_ms / 1100 * 1.05
    for name in ("f", "g"):
        assert np.array_equal(sims[0].populations(name), sims[1].populations(name))
    # a second timed call counts its own launches only
    sims[0].step(7)
    assert sims[0].timing().launches == 7
    sims[0].set_kernel_timing(False)
    sims[0].step(3)
    for s in sims:
        s.close()
