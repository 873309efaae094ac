"""CPU checks of the C-ABI boundary: the library loads, exports every entry point that
include/rbtp.h declares, and its host-side logic (case defaults, derived constants, the
benchmark initial temperature, argument validation) matches the oracle without touching a GPU.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from cases import config1, config2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "rbtp.h")).read()
    return sorted(set(re.findall(r"\b(rbtp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(rb):
    lib = ctypes.CDLL(rb.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_abi_version(rb):
    assert rb.lib.rbtp_abi_version() == 1


def test_case_struct_layout_matches_header(rb, O):
    assert ctypes.sizeof(rb.Case) == ctypes.sizeof(O.Case)
    assert [f[0] for f in rb.Case._fields_] == [f[0] for f in O.Case._fields_]


def test_case_defaults_are_parameters_h(rb, O):
    a = rb.default_case()
    b = O.default_case()
    assert bytes(a) == bytes(b)
    assert (a.lx, a.ly, a.lz, a.nend, a.ndiag) == (320, 320, 160, 1000000, 125)
    assert a.rayleigh == 1.0e7 and a.visc == 0.01206 and a.particle_rand_seed == 86143


@pytest.mark.parametrize("kw", [config1(), config2(), dict(particle_fraction=0.0096,
                                                            particle_diam_frac=1 / 32)])
def test_derived_constants_match_oracle(rb, O, kw):
    a = rb.derive(rb.default_case(**kw))
    b = O.derive(O.default_case(**kw))
    for k, _ in rb.Derived._fields_:
        if k != "reserved":
            assert getattr(a, k) == getattr(b, k), k


def test_linear_temperature_matches_oracle(rb, O):
    c = rb.default_case(**config1())
    t1 = rb.linear_temperature(c, 99, 1e-3)
    t2 = O.linear_temperature(c, 99, 1e-3)
    assert np.array_equal(t1, t2)
    x = (np.arange(t1.size) % 64 + 0.5) / 64
    assert np.max(np.abs(t1 - (0.5 - x))) <= 1e-3


def test_invalid_case_is_rejected_before_the_gpu(rb):
    c = rb.default_case(lx=2)
    h = ctypes.c_void_p()
    rc = rb.lib.rbtp_create(ctypes.byref(c), ctypes.byref(h))
    assert rc == -1 and not h.value
    assert b"extent" in rb.lib.rbtp_last_error()
    c = rb.default_case(precision=9)
    assert rb.lib.rbtp_create(ctypes.byref(c), ctypes.byref(h)) == -1


def test_null_arguments_return_errors(rb):
    assert rb.lib.rbtp_case_default(None) == -1
    assert rb.lib.rbtp_step(None, 1, 0) == -1
    assert rb.lib.rbtp_get_field(None, 0, np.zeros(1)) == -1
