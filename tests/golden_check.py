"""Replay of the golden fixtures (tests/golden/*.npz) on the oracle or on the HIP library.

A fixture holds a case's inputs and the oracle's state after a few steps (see
tests/golden/make_fixtures.py); replay() runs the same inputs through a solver with the Oracle
or Sim interface and compare() checks the result against the stored state.
"""
import glob
import hashlib
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(os.path.splitext(os.path.basename(p))[0]
                  for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def replay(mod, fx):
    """Run fixture fx's inputs on module mod (the oracle or rbtp_amd); returns the solver."""
    kw = json.loads(str(fx["case_json"]))
    sim = (mod.Oracle if hasattr(mod, "Oracle") else mod.Sim)(mod.default_case(**kw))
    pp = fx["ppos"] if fx["ppos"].size else None
    pt = fx["ptemp"] if fx["ptemp"].size else None
    sim.init(temperature=fx["temperature"], ppos=pp, ptemp=pt)
    if fx["pvel"].size:
        for a, comp in enumerate(("vel_x", "vel_y", "vel_z")):
            sim.set_particles(comp, np.ascontiguousarray(fx["pvel"][:, a]))
    sim.step(int(fx["steps"]))
    return sim


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def rel_err(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    if a.size == 0 and b.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def compare(sim, fx, field_tol=0.0, nu_tol=0.0):
    """Populations, flags, links and new-fluid lists bitwise; fields and particle arrays within
    field_tol relative and Nu within nu_tol (0: bitwise)."""
    n = sim.n
    for name in ("f", "g"):
        pop = sim.populations(name).reshape(27, n)
        if name in fx:
            assert np.array_equal(pop, fx[name]), f"{name}: max rel diff {rel_err(pop, fx[name]):.3e}"
        if "link_nodes" in fx:
            at = pop[:, fx["link_nodes"]]
            ref = fx[name + "_at_link_nodes"]
            assert np.array_equal(at, ref), f"{name} at link nodes: max rel diff {rel_err(at, ref):.3e}"
        assert digest(pop) == str(fx[name + "_sha256"]), f"{name}: digest differs"
    for name in ("rho", "ux", "uy", "uz", "phi"):
        v = sim.field(name)
        assert rel_err(v, fx["field_" + name]) <= field_tol, name
    nu = np.asarray(sim.nu())
    assert np.all(np.abs(nu - fx["nu"]) <= nu_tol * np.maximum(1.0, np.abs(fx["nu"]))), (nu, fx["nu"])
    if "ibnode" not in fx:
        return
    ib, ow = sim.flags()
    assert np.array_equal(np.asarray(ib), fx["ibnode"]) and np.array_equal(np.asarray(ow), fx["owner"])
    links = sim.links()
    order = np.lexsort((links["dir"], (links["cell_k"].astype(np.int64) * 100000 +
                                       links["cell_j"]) * 100000 + links["cell_i"]))
    links = links[order]
    assert len(links) == len(fx["link_q"])
    for k in ("cell_i", "cell_j", "cell_k", "dir", "pid"):
        assert np.array_equal(links[k], fx["link_" + k]), k
    for k in ("q", "rx", "ry", "rz"):
        assert np.array_equal(np.asarray(links[k]).view(np.uint64),
                              fx["link_" + k].view(np.uint64)), k
    nodes, pids = sim.new_fluid()
    assert np.array_equal(np.asarray(nodes), fx["new_fluid_nodes"])
    assert np.array_equal(np.asarray(pids), fx["new_fluid_pids"])
    for key in fx:
        if key.startswith("p_"):
            v = sim.particles(key[2:])
            assert rel_err(v, fx[key]) <= field_tol, key
