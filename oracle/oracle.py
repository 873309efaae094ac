"""ctypes view of the FP64 host oracle (oracle/liboracle.so).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module; the
product library never does.  The oracle restates the reference kernels on the CPU with the same
operation order (see oracle/oracle.cpp).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# RBTP_ORACLE_LIB names another build of oracle.cpp: tests/test_oracle_mutants.py points it at
# builds with a planted mistake (ORACLE_MUTANT) to show that the pins catch it.
LIB_PATH = os.environ.get("RBTP_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
ROOT = os.path.dirname(HERE)


# Case/Link layouts are plain data; define them here again so the oracle does not need the GPU
# library to be loadable.
class Case(C.Structure):
    _fields_ = [
        ("continuous_steps", C.c_int), ("lx", C.c_int), ("ly", C.c_int), ("lz", C.c_int),
        ("nend", C.c_int), ("ndiag", C.c_int), ("nflowout", C.c_int), ("nnuout", C.c_int),
        ("nout_p", C.c_int),
        ("rayleigh", C.c_double), ("prandtl", C.c_double), ("t_hot", C.c_double),
        ("t_cold", C.c_double), ("t_ref", C.c_double), ("beta", C.c_double), ("visc", C.c_double),
        ("rho0", C.c_double),
        ("activate_particles", C.c_int),
        ("particle_fraction", C.c_double), ("particle_diam_frac", C.c_double),
        ("cp_particle_ratio", C.c_double), ("cp_fluid", C.c_double),
        ("rho_particle_ratio", C.c_double), ("repforce_scale", C.c_double),
        ("particle_init_temp", C.c_double),
        ("particle_rand_seed", C.c_int),
        ("rep_min_gap", C.c_double), ("rep_min_gap_w", C.c_double), ("rep_stf0", C.c_double),
        ("rep_stf1", C.c_double), ("rep_stf0_w", C.c_double), ("rep_stf1_w", C.c_double),
        ("npart_override", C.c_int), ("precision", C.c_int), ("reserved", C.c_int * 6),
    ]

    def update(self, **kw):
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError(k)
            setattr(self, k, v)
        return self


class Derived(C.Structure):
    _fields_ = [
        ("grav0", C.c_double), ("tau", C.c_double), ("u0", C.c_double), ("diff", C.c_double),
        ("tauc", C.c_double), ("tauci", C.c_double), ("particle_radius", C.c_double),
        ("particle_mass", C.c_double), ("particle_inertia", C.c_double),
        ("particle_heat_capacity", C.c_double), ("npart", C.c_int), ("reserved", C.c_int),
    ]


LINK_DTYPE = np.dtype([
    ("cell_i", "<i4"), ("cell_j", "<i4"), ("cell_k", "<i4"), ("dir", "<i4"), ("pid", "<i4"),
    ("pad", "<i4"), ("q", "<f8"), ("rx", "<f8"), ("ry", "<f8"), ("rz", "<f8"),
])

# Stage ids of or_stage, in main.cu order.
STAGES = ["macrovar", "forcing", "collide_f", "stream_f", "ibb_f", "collide_g", "stream_g",
          "ibb_g", "prepare", "repulsion", "forces", "heat", "move", "build_links", "refill"]
PARTICLE_ARRAYS = [
    "pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "omega_x", "omega_y", "omega_z",
    "theta_x", "theta_y", "theta_z", "force_x", "force_y", "force_z", "torque_x", "torque_y",
    "torque_z", "force_prev_x", "force_prev_y", "force_prev_z", "force_prev2_x",
    "force_prev2_y", "force_prev2_z", "torque_prev_x", "torque_prev_y", "torque_prev_z",
    "torque_prev2_x", "torque_prev2_y", "torque_prev2_z", "temp", "heat", "heat_prev",
]
FIELD = {"rho": 0, "ux": 1, "uy": 2, "uz": 3, "phi": 4, "fx": 5, "fy": 6, "fz": 7}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
    lib = C.CDLL(LIB_PATH)
    P, I, D, L = C.c_void_p, C.c_int, C.c_double, C.c_long
    sig = {
        "or_abi_version": (I, []),
        "or_case_default": (None, [C.POINTER(Case)]),
        "or_case_derive": (None, [C.POINTER(Case), C.POINTER(Derived)]),
        "or_create": (P, [C.POINTER(Case)]),
        "or_destroy": (None, [P]),
        "or_init": (I, [P, P, P, P]),
        "or_stage": (I, [P, I]),
        "or_step": (I, [P, I]),
        "or_npart": (I, [P]),
        "or_nodes": (L, [P]),
        "or_current_step": (I, [P]),
        "or_get_derived": (None, [P, C.POINTER(Derived)]),
        "or_unhandled_links": (L, [P, I]),
        "or_array": (C.POINTER(C.c_double), [P, I]),
        "or_iarray": (C.POINTER(C.c_int), [P, I]),
        "or_nlinks": (L, [P]),
        "or_links": (P, [P]),
        "or_nnew_fluid": (L, [P]),
        "or_new_fluid_nodes": (C.POINTER(C.c_int), [P]),
        "or_new_fluid_pids": (C.POINTER(C.c_int), [P]),
        "or_linear_temperature": (None, [C.POINTER(Case), C.c_uint64, D, P]),
        "or_nu": (None, [P, P]),
        "or_means": (None, [P, P]),
        "or_threads": (I, []),
        "or_output_fields": (None, [P, P]),
        "or_profile": (None, [P, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def default_case(**kw) -> Case:
    c = Case()
    lib.or_case_default(C.byref(c))
    return c.update(**kw)


def derive(case) -> Derived:
    d = Derived()
    lib.or_case_derive(C.cast(C.byref(case), C.POINTER(Case)), C.byref(d))
    return d


def as_case(case) -> Case:
    """Copy any rbtp_case-layout ctypes structure into the oracle's Case."""
    return Case.from_buffer_copy(bytes(case))


def linear_temperature(case, seed: int, amp: float) -> np.ndarray:
    c = as_case(case)
    out = np.empty(c.lx * c.ly * c.lz, dtype=np.float64)
    lib.or_linear_temperature(C.byref(c), C.c_uint64(seed), amp, out.ctypes.data_as(C.c_void_p))
    return out


def threads() -> int:
    return lib.or_threads()


class Oracle:
    """FP64 host restatement of the reference's state and step sequence."""

    def __init__(self, case):
        self.case = as_case(case)
        self._h = lib.or_create(C.byref(self.case))
        if not self._h:
            raise ValueError("invalid case for the oracle")
        self.n = int(lib.or_nodes(self._h))
        self.np = int(lib.or_npart(self._h))
        self.derived = Derived()
        lib.or_get_derived(self._h, C.byref(self.derived))

    def close(self):
        if getattr(self, "_h", None):
            lib.or_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def init(self, temperature=None, ppos=None, ptemp=None) -> None:
        keep = []

        def ptr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        rc = lib.or_init(self._h, ptr(temperature), ptr(ppos), ptr(ptemp))
        if rc != 0:
            raise RuntimeError(f"oracle init failed ({rc})")

    def step(self, n: int = 1) -> None:
        lib.or_step(self._h, int(n))

    def stage(self, name: str) -> None:
        lib.or_stage(self._h, STAGES.index(name))

    @property
    def current_step(self) -> int:
        return int(lib.or_current_step(self._h))

    def _view(self, which: int, count: int) -> np.ndarray:
        p = lib.or_array(self._h, which)
        return np.ctypeslib.as_array(p, shape=(count,)).copy()

    def populations(self, name: str, copy: bool = True) -> np.ndarray:
        """27 x n populations; copy=False gives a view of the oracle's own memory (valid while
        the oracle lives and until its next stage), for lattices too large to copy."""
        which = {"f": 0, "g": 1, "f_collide": 2, "g_collide": 3}[name]
        if not copy:
            p = lib.or_array(self._h, which)
            return np.ctypeslib.as_array(p, shape=(27 * self.n,)).reshape(27, self.n)
        return self._view(which, 27 * self.n).reshape(27, self.n)

    def field(self, name: str) -> np.ndarray:
        return self._view(10 + FIELD[name], self.n)

    def set_field(self, name: str, values) -> None:
        p = lib.or_array(self._h, 10 + FIELD[name])
        np.ctypeslib.as_array(p, shape=(self.n,))[:] = values

    def set_populations(self, name: str, values) -> None:
        which = {"f": 0, "g": 1, "f_collide": 2, "g_collide": 3}[name]
        p = lib.or_array(self._h, which)
        np.ctypeslib.as_array(p, shape=(27 * self.n,))[:] = np.asarray(values).reshape(-1)

    def particles(self, name: str) -> np.ndarray:
        if self.np == 0:
            return np.zeros(0)
        return self._view(100 + PARTICLE_ARRAYS.index(name), self.np)

    def set_particles(self, name: str, values) -> None:
        p = lib.or_array(self._h, 100 + PARTICLE_ARRAYS.index(name))
        np.ctypeslib.as_array(p, shape=(self.np,))[:] = values

    def flags(self, prev: bool = False):
        ib = lib.or_iarray(self._h, 2 if prev else 0)
        ow = lib.or_iarray(self._h, 3 if prev else 1)
        return (np.ctypeslib.as_array(ib, shape=(self.n,)).copy(),
                np.ctypeslib.as_array(ow, shape=(self.n,)).copy())

    def links(self) -> np.ndarray:
        n = int(lib.or_nlinks(self._h))
        if n == 0:
            return np.zeros(0, dtype=LINK_DTYPE)
        buf = (C.c_char * (n * LINK_DTYPE.itemsize)).from_address(lib.or_links(self._h))
        return np.frombuffer(bytes(buf), dtype=LINK_DTYPE).copy()

    def new_fluid(self):
        n = int(lib.or_nnew_fluid(self._h))
        if n == 0:
            return np.zeros(0, np.int32), np.zeros(0, np.int32)
        nodes = np.ctypeslib.as_array(lib.or_new_fluid_nodes(self._h), shape=(n,)).copy()
        pids = np.ctypeslib.as_array(lib.or_new_fluid_pids(self._h), shape=(n,)).copy()
        return nodes.astype(np.int32), pids.astype(np.int32)

    def unhandled_links(self, thermal: bool) -> int:
        return int(lib.or_unhandled_links(self._h, 1 if thermal else 0))

    def nu(self) -> np.ndarray:
        out = np.zeros(7)
        lib.or_nu(self._h, out.ctypes.data_as(C.c_void_p))
        return out

    def output_fields(self) -> np.ndarray:
        out = np.zeros(7 * self.n)
        lib.or_output_fields(self._h, out.ctypes.data_as(C.c_void_p))
        return out.reshape(7, self.n)

    def profile(self) -> np.ndarray:
        fields = np.ascontiguousarray(self.output_fields().reshape(-1))
        out = np.zeros(4 * self.case.lx)
        lib.or_profile(self._h, fields.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
        return out.reshape(self.case.lx, 4)

    def means(self) -> np.ndarray:
        out = np.zeros(4)
        lib.or_means(self._h, out.ctypes.data_as(C.c_void_p))
        return out
