"""Python view of librbtp (include/rbtp.h) for tests, benchmarks and scripting.

The solver itself is the HIP library ``librbtp.so`` built next to this file; this module only
binds its C-ABI with ctypes and returns numpy arrays.  Importing it fails loudly when the library
has not been built: there is no Python or CPU fallback.

The stage names follow the reference's host API (include/particle.h:110-122,
include/lbm.h:52-84): ``Sim.step`` runs the main.cu:32-86 sequence, the accessors expose the
reference's arrays (d_f, d_g, d_f_collide, d_g_collide, rho/ux/uy/uz/phi, force_real*,
ibnode/owner, the particle SoA, the link list and the new-fluid list).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librbtp.so")

PREC_FP64, PREC_FP32, PREC_FP16S, PREC_BF16S = 0, 1, 2, 3
PREC_NAMES = {PREC_FP64: "fp64", PREC_FP32: "fp32", PREC_FP16S: "fp16s", PREC_BF16S: "bf16s"}
STEP_MACRO_LAST = 1

FIELD = {"rho": 0, "ux": 1, "uy": 2, "uz": 3, "phi": 4, "fx": 5, "fy": 6, "fz": 7}
POP = {"f": 0, "g": 1, "f_collide": 2, "g_collide": 3}
# RBTP_STAGE_*: the stages of one time step in main.cu order (main.cu:35-84).
STAGES = ["macrovar", "forcing", "collide_f", "stream_f", "ibb_f", "collide_g", "stream_g",
          "ibb_g", "prepare", "repulsion", "forces", "heat", "move", "build_links", "refill"]
PARTICLE_ARRAYS = [
    "pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "omega_x", "omega_y", "omega_z",
    "theta_x", "theta_y", "theta_z", "force_x", "force_y", "force_z", "torque_x", "torque_y",
    "torque_z", "force_prev_x", "force_prev_y", "force_prev_z", "force_prev2_x",
    "force_prev2_y", "force_prev2_z", "torque_prev_x", "torque_prev_y", "torque_prev_z",
    "torque_prev2_x", "torque_prev2_y", "torque_prev2_z", "temp", "heat", "heat_prev",
]

_CASE_FIELDS = [
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


class Case(C.Structure):
    """rbtp_case: every parameters.h macro as a runtime field (lower case) plus extensions."""

    _fields_ = _CASE_FIELDS

    def update(self, **kw) -> "Case":
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError(f"rbtp_case has no field {k!r}")
            setattr(self, k, v)
        return self

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in _CASE_FIELDS if k != "reserved"}


class Derived(C.Structure):
    _fields_ = [
        ("grav0", C.c_double), ("tau", C.c_double), ("u0", C.c_double), ("diff", C.c_double),
        ("tauc", C.c_double), ("tauci", C.c_double), ("particle_radius", C.c_double),
        ("particle_mass", C.c_double), ("particle_inertia", C.c_double),
        ("particle_heat_capacity", C.c_double), ("npart", C.c_int), ("reserved", C.c_int),
    ]


class Link(C.Structure):
    """ParticleLink (include/particle.h:19-29), 56 bytes."""

    _fields_ = [
        ("cell_i", C.c_int32), ("cell_j", C.c_int32), ("cell_k", C.c_int32), ("dir", C.c_int32),
        ("pid", C.c_int32), ("pad", C.c_int32), ("q", C.c_double), ("rx", C.c_double),
        ("ry", C.c_double), ("rz", C.c_double),
    ]


LINK_DTYPE = np.dtype([
    ("cell_i", "<i4"), ("cell_j", "<i4"), ("cell_k", "<i4"), ("dir", "<i4"), ("pid", "<i4"),
    ("pad", "<i4"), ("q", "<f8"), ("rx", "<f8"), ("ry", "<f8"), ("rz", "<f8"),
])
assert LINK_DTYPE.itemsize == C.sizeof(Link) == 56


class RbtpError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rbtp error {code}: {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C rb-thermoparticles_amd` "
            "(or __graft_entry__.build()); there is no fallback implementation")
    lib = C.CDLL(LIB_PATH)
    P, I, D, U = C.c_void_p, C.c_int, C.c_double, C.c_uint
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    ip32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    sig = {
        "rbtp_last_error": (C.c_char_p, []),
        "rbtp_abi_version": (I, []),
        "rbtp_case_default": (I, [C.POINTER(Case)]),
        "rbtp_case_derive": (I, [C.POINTER(Case), C.POINTER(Derived)]),
        "rbtp_create": (I, [C.POINTER(Case), C.POINTER(P)]),
        "rbtp_destroy": (None, [P]),
        "rbtp_get_case": (I, [P, C.POINTER(Case), C.POINTER(Derived)]),
        "rbtp_init": (I, [P, P, P, P]),
        "rbtp_linear_temperature": (I, [C.POINTER(Case), C.c_uint64, D, dp]),
        "rbtp_step": (I, [P, I, U]),
        "rbtp_step_async": (I, [P, I, U]),
        "rbtp_synchronize": (I, [P]),
        "rbtp_current_step": (I, [P, C.POINTER(I)]),
        "rbtp_stage": (I, [P, I]),
        "rbtp_next_stage": (I, [P, C.POINTER(I)]),
        "rbtp_get_field": (I, [P, I, dp]),
        "rbtp_get_populations": (I, [P, I, dp]),
        "rbtp_set_populations": (I, [P, dp, dp]),
        "rbtp_get_flags": (I, [P, ip32, ip32]),
        "rbtp_get_particles": (I, [P, I, dp]),
        "rbtp_set_particles": (I, [P, I, dp]),
        "rbtp_get_links": (I, [P, P, C.c_int64, C.POINTER(C.c_int64)]),
        "rbtp_get_new_fluid": (I, [P, P, P, C.c_int64, C.POINTER(C.c_int64)]),
        "rbtp_nu": (I, [P, dp]),
        "rbtp_nu_noalloc": (I, [P, dp]),
        "rbtp_means": (I, [P, dp]),
        "rbtp_kinetic_energy": (I, [P, C.POINTER(D)]),
        "rbtp_case_from_parameters_h": (I, [C.c_char_p, C.POINTER(Case)]),
        "rbtp_bandwidth_probe": (I, [C.c_size_t, I, C.POINTER(D)]),
        "rbtp_stencil_probe": (I, [I, I, I, I, C.POINTER(D)]),
        "rbtp_get_output_fields": (I, [P, dp]),
        "rbtp_profile": (I, [P, dp]),
        "rbtp_write_fg": (I, [P, C.c_char_p]),
        "rbtp_write_flow": (I, [P, C.c_char_p]),
        "rbtp_write_particles": (I, [P, C.c_char_p]),
        "rbtp_restart_load": (I, [P, C.c_char_p, I]),
        "rbtp_batch_create": (I, [I, I, D, C.POINTER(P)]),
        "rbtp_batch_destroy": (None, [P]),
        "rbtp_batch_admit": (I, [P, C.POINTER(Case), P, P, P, C.POINTER(I)]),
        "rbtp_batch_release": (I, [P, I]),
        "rbtp_batch_slot": (I, [P, I, C.POINTER(P)]),
        "rbtp_batch_step": (I, [P, I, U]),
        "rbtp_batch_tune": (I, [P, C.POINTER(I), C.POINTER(I), C.POINTER(D)]),
        "rbtp_batch_set_launch": (I, [P, I, I, I]),
        "rbtp_batch_info": (I, [P, C.POINTER(I), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                C.POINTER(D), C.POINTER(I), C.POINTER(I)]),
        "rbtp_last_step_ms": (I, [P, C.POINTER(D)]),
        "rbtp_set_kernel_timing": (I, [P, I]),
        "rbtp_kernel_ms": (I, [P, C.POINTER(D), C.POINTER(C.c_int64)]),
        "rbtp_set_shift_interval": (I, [P, I]),
        "rbtp_set_watchdog": (I, [P, D]),
        "rbtp_streaming_probe": (I, [I, I, I, I, I, C.POINTER(D)]),
        "rbtp_arith_probe": (I, [I, dp, P, C.c_int64, dp]),
        "rbtp_place_particles": (I, [C.POINTER(Case), dp]),
        "rbtp_read_particle_file": (I, [C.c_char_p, I, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _check(rc: int) -> None:
    if rc != 0:
        raise RbtpError(rc, lib.rbtp_last_error().decode(errors="replace"))


def default_case(**kw) -> Case:
    c = Case()
    _check(lib.rbtp_case_default(C.byref(c)))
    return c.update(**kw)


def case_from_parameters_h(path: str) -> Case:
    """Case from a parameters.h file (same macro names as the reference)."""
    c = Case()
    _check(lib.rbtp_case_from_parameters_h(os.fsencode(path), C.byref(c)))
    return c


def bandwidth_probe(nbytes: int = 2 << 30, reps: int = 20) -> float:
    """Achievable HBM bandwidth (GB/s) of a streaming copy on the current device."""
    v = C.c_double()
    _check(lib.rbtp_bandwidth_probe(int(nbytes), int(reps), C.byref(v)))
    return v.value


def streaming_probe(scheme: int, lx: int, ly: int, lz: int, reps: int = 20) -> float:
    """GB/s of a streaming scheme's access pattern (rbtp_streaming_probe).  FP64, at 880 B per
    node: 0 the pull of the FP64 kernel, 1 and 2 the even and odd steps of in-place AA
    streaming, 3 the plain pull.  16-bit storage, at 224 B per node: 4 the plain pull of the
    FP16/BF16 kernels, 5 the row pull with lane shifts, 6 two nodes per thread."""
    v = C.c_double()
    _check(lib.rbtp_streaming_probe(int(scheme), int(lx), int(ly), int(lz), int(reps), C.byref(v)))
    return v.value


def stencil_probe(lx: int, ly: int, lz: int, reps: int = 20) -> float:
    """GB/s of the FP64 lattice step's access pattern alone (rbtp_stencil_probe), at 880 B per
    node: the ceiling of the pattern on this device."""
    v = C.c_double()
    _check(lib.rbtp_stencil_probe(int(lx), int(ly), int(lz), int(reps), C.byref(v)))
    return v.value


ARITH = {"div": 0, "div_hostrecip": 1, "div6": 2, "div_rt": 3, "sqrt": 4}


def arith_probe(op: str, a, b=None) -> np.ndarray:
    """The strict kernels' FP64 division or square root evaluated on the device
    (rbtp_arith_probe): op is one of ARITH; b is the divisor for "div" and "div_hostrecip"."""
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    bp = None
    if b is not None:
        b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1)
        if b.size != a.size:
            raise ValueError("a and b must have the same length")
        bp = b.ctypes.data_as(C.c_void_p)
    out = np.empty_like(a)
    _check(lib.rbtp_arith_probe(ARITH[op], a, bp, a.size, out))
    return out


def place_particles(case: Case) -> np.ndarray:
    """Reference rejection placement (pinit.cu:97-144) on the host: shape (3, npart)."""
    npart = derive(case).npart
    out = np.zeros(3 * max(npart, 0), dtype=np.float64)
    if npart > 0:
        _check(lib.rbtp_place_particles(C.byref(case), out))
    return out.reshape(3, max(npart, 0))


def read_particle_file(path: str, npart: int) -> dict:
    """A pos_%09d.txt file (initial.cu:57-192) as a dict of particle arrays (host only)."""
    out = np.zeros(len(PARTICLE_ARRAYS) * max(npart, 0), dtype=np.float64)
    _check(lib.rbtp_read_particle_file(os.fsencode(path), int(npart), out))
    out = out.reshape(len(PARTICLE_ARRAYS), max(npart, 0))
    return {name: out[k] for k, name in enumerate(PARTICLE_ARRAYS)}


def derive(case: Case) -> Derived:
    d = Derived()
    _check(lib.rbtp_case_derive(C.byref(case), C.byref(d)))
    return d


def linear_temperature(case: Case, seed: int, amp: float) -> np.ndarray:
    out = np.empty(case.lx * case.ly * case.lz, dtype=np.float64)
    _check(lib.rbtp_linear_temperature(C.byref(case), C.c_uint64(seed), amp, out))
    return out


@dataclass
class StepTiming:
    step_ms: float
    kernel_ms: float
    launches: int


class Sim:
    """One solver instance (rbtp_sim) on the current HIP device."""

    def __init__(self, case: Case, _handle=None):
        self.case = Case.from_buffer_copy(case)
        self._owned = _handle is None
        if _handle is None:
            h = C.c_void_p()
            _check(lib.rbtp_create(C.byref(self.case), C.byref(h)))
        else:
            h = _handle
        self._h = h
        self.derived = Derived()
        _check(lib.rbtp_get_case(self._h, None, C.byref(self.derived)))
        self.n = case.lx * case.ly * case.lz
        self.np = self.derived.npart if case.activate_particles else 0

    def close(self) -> None:
        if getattr(self, "_h", None):
            if self._owned:
                lib.rbtp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- initial state and stepping ------------------------------------------------------------
    def init(self, temperature=None, ppos=None, ptemp=None) -> None:
        t = None if temperature is None else np.ascontiguousarray(temperature, dtype=np.float64)
        p = None if ppos is None else np.ascontiguousarray(ppos, dtype=np.float64).reshape(-1)
        pt = None if ptemp is None else np.ascontiguousarray(ptemp, dtype=np.float64)
        if t is not None and t.size != self.n:
            raise ValueError("temperature must hold lx*ly*lz values")
        if p is not None and p.size != 3 * self.np:
            raise ValueError("ppos must hold 3*npart values (x[], y[], z[])")
        if pt is not None and pt.size != self.np:
            raise ValueError("ptemp must hold npart values")
        self._keep = (t, p, pt)
        ptr = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        _check(lib.rbtp_init(self._h, ptr(t), ptr(p), ptr(pt)))

    def step(self, nsteps: int = 1, macro_last: bool = True) -> None:
        _check(lib.rbtp_step(self._h, int(nsteps), STEP_MACRO_LAST if macro_last else 0))

    def step_async(self, nsteps: int = 1, macro_last: bool = False) -> None:
        _check(lib.rbtp_step_async(self._h, int(nsteps), STEP_MACRO_LAST if macro_last else 0))

    def synchronize(self) -> None:
        _check(lib.rbtp_synchronize(self._h))

    def stage(self, name: str) -> None:
        """Run one stage of the time step (rbtp_stage), in main.cu order; FP64 form only."""
        _check(lib.rbtp_stage(self._h, STAGES.index(name)))

    @property
    def next_stage(self) -> str:
        v = C.c_int()
        _check(lib.rbtp_next_stage(self._h, C.byref(v)))
        return STAGES[v.value]

    @property
    def current_step(self) -> int:
        v = C.c_int()
        _check(lib.rbtp_current_step(self._h, C.byref(v)))
        return v.value

    # -- accessors --------------------------------------------------------------------------------
    def field(self, name: str) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float64)
        _check(lib.rbtp_get_field(self._h, FIELD[name], out))
        return out

    def populations(self, name: str) -> np.ndarray:
        out = np.empty(27 * self.n, dtype=np.float64)
        _check(lib.rbtp_get_populations(self._h, POP[name], out))
        return out.reshape(27, self.n)

    def set_populations(self, f: np.ndarray, g: np.ndarray) -> None:
        f = np.ascontiguousarray(f, np.float64).reshape(-1)
        g = np.ascontiguousarray(g, np.float64).reshape(-1)
        if f.size != 27 * self.n or g.size != 27 * self.n:
            raise ValueError(f"f and g must hold 27*n = {27 * self.n} values each "
                             f"(got {f.size} and {g.size})")
        _check(lib.rbtp_set_populations(self._h, f, g))

    def flags(self):
        ib = np.empty(self.n, dtype=np.int32)
        ow = np.empty(self.n, dtype=np.int32)
        _check(lib.rbtp_get_flags(self._h, ib, ow))
        return ib, ow

    def particles(self, name: str) -> np.ndarray:
        out = np.zeros(max(self.np, 0), dtype=np.float64)
        if self.np:
            _check(lib.rbtp_get_particles(self._h, PARTICLE_ARRAYS.index(name), out))
        return out

    def set_particles(self, name: str, values) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if v.size != self.np:
            raise ValueError(f"{name} must hold npart = {self.np} values (got {v.size})")
        if self.np:
            _check(lib.rbtp_set_particles(self._h, PARTICLE_ARRAYS.index(name), v))

    def links(self) -> np.ndarray:
        cnt = C.c_int64()
        _check(lib.rbtp_get_links(self._h, None, 0, C.byref(cnt)))
        out = np.zeros(cnt.value, dtype=LINK_DTYPE)
        if cnt.value:
            _check(lib.rbtp_get_links(self._h, out.ctypes.data_as(C.c_void_p), cnt.value,
                                      C.byref(cnt)))
        return out

    def new_fluid(self):
        cnt = C.c_int64()
        _check(lib.rbtp_get_new_fluid(self._h, None, None, 0, C.byref(cnt)))
        nodes = np.zeros(cnt.value, dtype=np.int32)
        pids = np.zeros(cnt.value, dtype=np.int32)
        if cnt.value:
            _check(lib.rbtp_get_new_fluid(self._h, nodes.ctypes.data_as(C.c_void_p),
                                          pids.ctypes.data_as(C.c_void_p), cnt.value,
                                          C.byref(cnt)))
        return nodes, pids

    # -- diagnostics -------------------------------------------------------------------------------
    def nu(self) -> np.ndarray:
        out = np.zeros(7, dtype=np.float64)
        _check(lib.rbtp_nu(self._h, out))
        return out

    def nu_noalloc(self) -> np.ndarray:
        """rbtp_nu_noalloc: output_nu with the partial sums in the library's static device
        scratch (no allocation)."""
        out = np.zeros(7, dtype=np.float64)
        _check(lib.rbtp_nu_noalloc(self._h, out))
        return out

    def means(self) -> np.ndarray:
        out = np.zeros(4, dtype=np.float64)
        _check(lib.rbtp_means(self._h, out))
        return out

    def kinetic_energy(self) -> float:
        v = C.c_double()
        _check(lib.rbtp_kinetic_energy(self._h, C.byref(v)))
        return v.value

    def output_fields(self) -> np.ndarray:
        """The seven %09d.dat fields (rho, ux, uy, uz, phi, ted, Ked) after the particle-node
        overwrite, shape (7, n)."""
        out = np.empty(7 * self.n, dtype=np.float64)
        _check(lib.rbtp_get_output_fields(self._h, out))
        return out.reshape(7, self.n)

    def profile(self) -> np.ndarray:
        out = np.empty(4 * self.case.lx, dtype=np.float64)
        _check(lib.rbtp_profile(self._h, out))
        return out.reshape(self.case.lx, 4)

    # -- files (fg%09d.dat, %09d.dat, pos_%09d.txt) and restart ----------------------------
    def write_fg(self, path: str) -> None:
        _check(lib.rbtp_write_fg(self._h, os.fsencode(path)))

    def write_flow(self, path: str) -> None:
        _check(lib.rbtp_write_flow(self._h, os.fsencode(path)))

    def write_particles(self, path: str) -> None:
        _check(lib.rbtp_write_particles(self._h, os.fsencode(path)))

    def restart_load(self, directory: str, step: int) -> None:
        _check(lib.rbtp_restart_load(self._h, os.fsencode(directory), int(step)))

    # -- timing ------------------------------------------------------------------------------------
    def set_kernel_timing(self, enable: bool) -> None:
        _check(lib.rbtp_set_kernel_timing(self._h, 1 if enable else 0))

    def set_watchdog(self, umax: float) -> None:
        """Divergence watchdog (rbtp_set_watchdog): with umax > 0, step() raises RbtpError
        (code -8, RBTP_E_DIVERGED) after steps in which a node's density or velocity was not
        finite or |u| exceeded umax; 0 turns it off."""
        _check(lib.rbtp_set_watchdog(self._h, float(umax)))

    def set_shift_interval(self, steps: int) -> None:
        """Steps between refreshes of the reduced-precision shift profile (0: never)."""
        _check(lib.rbtp_set_shift_interval(self._h, int(steps)))

    def timing(self) -> StepTiming:
        ms = C.c_double()
        _check(lib.rbtp_last_step_ms(self._h, C.byref(ms)))
        kms = C.c_double()
        n = C.c_int64()
        _check(lib.rbtp_kernel_ms(self._h, C.byref(kms), C.byref(n)))
        return StepTiming(ms.value, kms.value, n.value)


class CapacityError(RbtpError):
    """No free slot or not enough free HBM for an admission."""


class Batch:
    """A serving batch (rbtp_batch): independent cavities resident in HBM, stepped together by
    one batched persistent lattice launch per step, replayed from a captured hipGraph."""

    def __init__(self, precision: int = PREC_FP16S, max_slots: int = 64,
                 hbm_fraction: float = 0.8):
        h = C.c_void_p()
        _check(lib.rbtp_batch_create(int(precision), int(max_slots), float(hbm_fraction),
                                     C.byref(h)))
        self._h = h
        self.precision = precision
        self._slots = {}

    def close(self) -> None:
        if getattr(self, "_h", None):
            for v in self._slots.values():
                v._h = None
            lib.rbtp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def admit(self, case: Case, temperature=None, ppos=None, ptemp=None) -> int:
        """Admit a case; returns its slot id.  Raises CapacityError when full."""
        case = Case.from_buffer_copy(case)
        case.precision = self.precision
        n = case.lx * case.ly * case.lz
        npart = derive(case).npart if case.activate_particles else 0
        keep = []

        def ptr(a, size, what):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            if a.size != size:
                raise ValueError(f"{what} must hold {size} values (got {a.size})")
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        slot = C.c_int(-1)
        rc = lib.rbtp_batch_admit(self._h, C.byref(case), ptr(temperature, n, "temperature"),
                                  ptr(ppos, 3 * npart, "ppos (x[], y[], z[])"),
                                  ptr(ptemp, npart, "ptemp"), C.byref(slot))
        if rc == -6:
            raise CapacityError(rc, lib.rbtp_last_error().decode(errors="replace"))
        _check(rc)
        h = C.c_void_p()
        _check(lib.rbtp_batch_slot(self._h, slot.value, C.byref(h)))
        self._slots[slot.value] = Sim(case, _handle=h)
        return slot.value

    def release(self, slot: int) -> None:
        v = self._slots.pop(slot)
        v._h = None
        _check(lib.rbtp_batch_release(self._h, int(slot)))

    def sim(self, slot: int) -> Sim:
        return self._slots[slot]

    @property
    def slots(self):
        return list(self._slots)

    def step(self, nsteps: int, macro_last: bool = True) -> None:
        _check(lib.rbtp_batch_step(self._h, int(nsteps), STEP_MACRO_LAST if macro_last else 0))

    def tune(self):
        b, g, ms = C.c_int(), C.c_int(), C.c_double()
        _check(lib.rbtp_batch_tune(self._h, C.byref(b), C.byref(g), C.byref(ms)))
        return b.value, g.value, ms.value

    def set_launch(self, block: int, grid: int, graph_steps: int) -> None:
        _check(lib.rbtp_batch_set_launch(self._h, int(block), int(grid), int(graph_steps)))

    def info(self) -> dict:
        a, u, f, ms, b, g = C.c_int(), C.c_int64(), C.c_int64(), C.c_double(), C.c_int(), C.c_int()
        _check(lib.rbtp_batch_info(self._h, C.byref(a), C.byref(u), C.byref(f), C.byref(ms),
                                   C.byref(b), C.byref(g)))
        return dict(active=a.value, used_bytes=u.value, free_bytes=f.value, last_ms=ms.value,
                    block=b.value, grid=g.value)
