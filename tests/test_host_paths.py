"""CPU tests of the reference's host-side failure paths, through the C-ABI without a GPU:
the placement give-up (pinit.cu:126-131) and the particle restart reader (initial.cu:57-192).
"""
import numpy as np
import pytest

from cases import particle_case

# Column order of pos_%09d.txt after the pid (output_part.cu:101-113).
POS_COLUMNS = [
    "pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "force_x", "force_y", "force_z",
    "force_prev_x", "force_prev_y", "force_prev_z", "force_prev2_x", "force_prev2_y",
    "force_prev2_z", "theta_x", "theta_y", "theta_z", "omega_x", "omega_y", "omega_z",
    "torque_x", "torque_y", "torque_z", "torque_prev_x", "torque_prev_y", "torque_prev_z",
    "torque_prev2_x", "torque_prev2_y", "torque_prev2_z", "temp", "heat", "heat_prev",
]


def test_placement_matches_oracle(rb, O):
    """rbtp_place_particles gives the positions the oracle's rbtp_init-equivalent places (same
    glibc srand/rand sequence, pinit.cu:97-144)."""
    kw = particle_case(40, 40, 40, npart=12, diam=6.0, ra=1e4, beta=0.05)
    pos = rb.place_particles(rb.default_case(**kw))
    ora = O.Oracle(O.default_case(**kw))
    ora.init()
    for a, name in enumerate(("pos_x", "pos_y", "pos_z")):
        assert np.array_equal(pos[a], ora.particles(name)), name
    R = 3.0
    assert np.all(pos >= 1.5 * R) and np.all(pos <= 40 - 1.5 * R)
    d = np.sqrt(((pos[:, :, None] - pos[:, None, :]) ** 2).sum(axis=0))
    d[np.arange(12), np.arange(12)] = np.inf
    assert np.min(d) - 2 * R >= R / 3.0


def test_placement_gives_up(rb, O):
    """40 spheres of diameter 6 cannot fit a 16^3 box with the R/3 clearance: the reference
    prints "Can't distribute the particles" and exits; the library returns RBTP_E_PLACEMENT."""
    kw = particle_case(16, 16, 16, npart=40, diam=6.0, ra=1e4, beta=0.05)
    with pytest.raises(rb.RbtpError) as exc:
        rb.place_particles(rb.default_case(**kw))
    assert exc.value.code == -7
    assert "Can't distribute the particles" in str(exc.value)
    ora = O.Oracle(O.default_case(**kw))
    with pytest.raises(RuntimeError):
        ora.init()


def write_pos(path, rows):
    with open(path, "w") as fh:
        for pid, vals in rows:
            fh.write("%d" % pid + "".join(" %14.6e" % v for v in vals) + "\n")


def test_particle_file_columns(rb, tmp_path):
    rng = np.random.default_rng(5)
    vals = rng.uniform(-10, 10, size=(3, 33))
    path = tmp_path / "pos_000000100.txt"
    # rows out of pid order: the reader places each line by its pid
    write_pos(path, [(2, vals[2]), (0, vals[0]), (1, vals[1])])
    got = rb.read_particle_file(str(path), 3)
    for c, name in enumerate(POS_COLUMNS):
        expect = np.array([float("%14.6e" % vals[p, c]) for p in range(3)])
        assert np.array_equal(got[name], expect), name


@pytest.mark.parametrize("bad", ["short", "token", "pid_high", "pid_negative", "lines"])
def test_malformed_particle_file(rb, tmp_path, bad):
    """A short line, an unparsable value, a pid outside [0, np) or too few lines is RBTP_E_IO
    (initial.cu:88-101 exits; the reference also indexes with an unchecked pid)."""
    path = tmp_path / "pos_000000100.txt"
    row = list(np.linspace(0.5, 3.3, 33))
    if bad == "short":
        with open(path, "w") as fh:
            fh.write("0" + "".join(" %14.6e" % v for v in row) + "\n")
            fh.write("1" + "".join(" %14.6e" % v for v in row[:20]) + "\n")
    elif bad == "token":
        with open(path, "w") as fh:
            fh.write("0" + "".join(" %14.6e" % v for v in row) + "\n")
            fh.write("1 abc" + "".join(" %14.6e" % v for v in row[1:]) + "\n")
    elif bad == "pid_high":
        write_pos(path, [(0, row), (2, row)])
    elif bad == "pid_negative":
        write_pos(path, [(0, row), (-1, row)])
    else:
        write_pos(path, [(0, row)])
    with pytest.raises(rb.RbtpError) as exc:
        rb.read_particle_file(str(path), 2)
    assert exc.value.code == -4
    assert "Failed to read particle restart data" in str(exc.value)


def test_missing_particle_file(rb, tmp_path):
    with pytest.raises(rb.RbtpError) as exc:
        rb.read_particle_file(str(tmp_path / "nope.txt"), 2)
    assert exc.value.code == -4
    assert "Failed to open restart particle file" in str(exc.value)
