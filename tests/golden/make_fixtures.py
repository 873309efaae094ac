"""Writes the golden fixtures of this directory from the FP64 oracle (oracle/, CPU only).

    python tests/golden/make_fixtures.py

Each fixture is a small .npz holding a case's inputs (case fields, the initial temperature, the
sphere positions and temperatures) and the oracle's state after a few steps: the post-stream f
and g (in full for the fluid case, as SHA-256 digests of their FP64 bytes plus the values at the
link nodes for the sphere cases), the macroscopic fields, output_nu's seven numbers, the
solid/fluid flags and owners of both geometry generations, the link records and the new-fluid
list, and every particle array.  tests/test_golden.py checks that the oracle still reproduces
them bit for bit (so a change of the oracle is seen), and tests/test_gpu_golden.py that the HIP
library does.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from cases import particle_case, single_phase  # noqa: E402

P_ARRAYS = ["pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "omega_x", "omega_y",
            "omega_z", "theta_x", "theta_y", "theta_z", "force_x", "force_y", "force_z",
            "torque_x", "torque_y", "torque_z", "force_prev_x", "force_prev_y", "force_prev_z",
            "force_prev2_x", "force_prev2_y", "force_prev2_z", "torque_prev_x", "torque_prev_y",
            "torque_prev_z", "torque_prev2_x", "torque_prev2_y", "torque_prev2_z", "temp",
            "heat", "heat_prev"]
LINK_FIELDS = ("cell_i", "cell_j", "cell_k", "dir", "pid", "q", "rx", "ry", "rz")

# name: (case fields, initial-temperature seed and amplitude, sphere positions, sphere
# temperatures, initial sphere velocities, steps, whether f and g are stored in full)
CASES = {
    "fluid_8x8x4": (single_phase(8, 8, 4, ra=1e5), (3, 1e-2), None, None, None, 10, True),
    "sphere_12x10x10": (particle_case(12, 10, 10, npart=1, diam=5.0, ra=1e5, beta=0.5),
                        (5, 1e-2), [(6.3, 5.2, 4.9)], [0.2], None, 10, False),
    "moving_spheres_16x12x12": (particle_case(16, 12, 12, npart=2, diam=4.0, ra=1e5, beta=0.5),
                                (7, 1e-2), [(5.2, 6.1, 5.8), (10.9, 5.7, 6.4)], [0.3, -0.1],
                                [(0.15, 0.02, 0.0), (-0.12, 0.0, 0.03)], 12, False),
}


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def state(sim, full_pop: bool) -> dict:
    """The observables a fixture pins, read through the Oracle or Sim accessors (same names)."""
    out = {}
    for name in ("f", "g"):
        pop = sim.populations(name).reshape(27, -1)
        out[name + "_sha256"] = digest(pop)
        if full_pop:
            out[name] = pop
    for name in ("rho", "ux", "uy", "uz", "phi"):
        out["field_" + name] = sim.field(name)
    out["nu"] = np.asarray(sim.nu())
    if sim.np:
        ib, ow = sim.flags()
        out["ibnode"], out["owner"] = np.asarray(ib), np.asarray(ow)
        links = sim.links()
        order = np.lexsort((links["dir"], (links["cell_k"].astype(np.int64) * 100000 +
                                           links["cell_j"]) * 100000 + links["cell_i"]))
        links = links[order]
        for k in LINK_FIELDS:
            out["link_" + k] = np.asarray(links[k])
        nodes, pids = sim.new_fluid()
        out["new_fluid_nodes"], out["new_fluid_pids"] = np.asarray(nodes), np.asarray(pids)
        for name in P_ARRAYS:
            out["p_" + name] = sim.particles(name)
        # populations at the link nodes, for a readable report when a digest differs
        n = sim.n
        lnodes = np.unique((links["cell_k"].astype(np.int64) * (sim.case.ly) + links["cell_j"]) *
                           sim.case.lx + links["cell_i"])
        out["link_nodes"] = lnodes
        out["f_at_link_nodes"] = sim.populations("f").reshape(27, n)[:, lnodes]
        out["g_at_link_nodes"] = sim.populations("g").reshape(27, n)[:, lnodes]
    return out


def run_case(make_sim, name):
    kw, (seed, amp), ppos, ptemp, pvel, steps, full_pop = CASES[name]
    sim, T = make_sim(kw, seed, amp)
    pp = None if ppos is None else np.asarray(ppos, float).T.reshape(-1)
    pt = None if ptemp is None else np.asarray(ptemp, float)
    sim.init(temperature=T, ppos=pp, ptemp=pt)
    if pvel is not None:
        v = np.asarray(pvel, float)
        for a, comp in enumerate(("vel_x", "vel_y", "vel_z")):
            sim.set_particles(comp, v[:, a].copy())
    sim.step(steps)
    inputs = dict(case_json=json.dumps(kw), temperature=T, steps=steps,
                  ppos=np.zeros(0) if pp is None else pp, ptemp=np.zeros(0) if pt is None else pt,
                  pvel=np.zeros(0) if pvel is None else np.asarray(pvel, float))
    return sim, inputs, state(sim, full_pop)


def main():
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    import oracle as O

    def make_oracle(kw, seed, amp):
        case = O.default_case(**kw)
        return O.Oracle(case), O.linear_temperature(case, seed, amp)

    for name in CASES:
        sim, inputs, st = run_case(make_oracle, name)
        sim.close()
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **inputs, **st)
        print(f"{path}: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
