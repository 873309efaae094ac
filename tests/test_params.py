"""CPU tests of the parameters.h reader (rbtp_case_from_parameters_h)."""
import textwrap

import pytest


def write(tmp_path, body):
    p = tmp_path / "parameters.h"
    p.write_text(textwrap.dedent(body))
    return str(p)


def test_reads_macros_with_c_arithmetic(rb, tmp_path):
    path = write(tmp_path, """\
        #ifndef CASE_H
        #define CASE_H
        // restart step
        #define CONTINUOUS_STEPS 0
        #define LX 48      /* x is the buoyancy direction */
        #define LY 40
        #define LZ (LX/2)
        #define LXY (LX * LY)
        #define LXYZ (LX * LY * LZ)
        #define NPOP 27
        #define NEND 500
        #define NDIAG 50
        #define RAYLEIGH 2.5e6
        #define PRANDTL 4.3
        #define T_HOT 0.5
        #define T_COLD -0.5
        #define T_REF 0.0
        #define VISC 0.004
        #define ACTIVATE_PARTICLES 1
        #define PARTICLE_DIAM_FRAC (1.0/12.0)
        #define PARTICLE_INIT_TEMP (T_COLD + T_HOT)/2
        #define PARTICLE_RAND_SEED 12345
        #define REP_STF1 0.004
        #define NPART_OVERRIDE 7
        #define BLOCK_X 8
        #endif
        """)
    c = rb.case_from_parameters_h(path)
    assert (c.lx, c.ly, c.lz) == (48, 40, 24)
    assert (c.nend, c.ndiag) == (500, 50)
    assert c.rayleigh == 2.5e6 and c.prandtl == 4.3
    assert c.t_cold == -0.5 and c.visc == 0.004
    assert c.particle_diam_frac == 1.0 / 12.0
    assert c.particle_init_temp == 0.0
    assert c.particle_rand_seed == 12345 and c.rep_stf1 == 0.004
    assert c.npart_override == 7
    # untouched fields keep the parameters.h defaults
    d = rb.default_case()
    assert c.beta == d.beta and c.nflowout == d.nflowout and c.rep_min_gap == d.rep_min_gap


def test_integer_division_truncates_like_c(rb, tmp_path):
    path = write(tmp_path, """\
        #define LX 50
        #define LZ (LX/3)
        #define PARTICLE_DIAM_FRAC (1/20)
        """)
    c = rb.case_from_parameters_h(path)
    assert c.lz == 16
    assert c.particle_diam_frac == 0.0


def test_errors_are_reported(rb, tmp_path):
    with pytest.raises(rb.RbtpError) as e:
        rb.case_from_parameters_h(write(tmp_path, "#define LX (NOT_DEFINED + 1)\n"))
    assert "NOT_DEFINED" in str(e.value)
    with pytest.raises(rb.RbtpError):
        rb.case_from_parameters_h(str(tmp_path / "missing.h"))
