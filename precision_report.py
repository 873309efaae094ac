"""Precision forms against FP64 on the production cases.

    python precision_report.py run --workload c5 --precision fp64 --steps 20000 --out c5_fp64.npz
    python precision_report.py run --workload c5 --precision fp16s --steps 20000 --out c5_fp16s.npz
    python precision_report.py compare c5_fp64.npz c5_fp32.npz c5_fp16s.npz

Workloads (synthetic initial state: linear temperature + 1e-3 seeded noise, fluid at rest,
spheres placed by the reference's glibc-rand rejection sampler, each sphere starting at the
fluid temperature of its centre cell):
  c1: 64x64x32, Ra=1e5, Pr=0.71, no spheres (laminar: the flow settles to a steady state);
  c2: 320x320x160, Ra=1e8, Pr=0.71, no spheres;
  c3: 320x320x160, Ra=1e8, Pr=0.71, 300 spheres of diameter 10;
  c5: 512x512x256, Ra=1e9, Pr=0.71, 5000 spheres of diameter 8.

--perturb EPS adds EPS times a second seeded noise field to the initial temperature.  An FP64
run with a tiny EPS (1e-12) against the plain FP64 run measures how far the observables of a
chaotic flow drift from a perturbation far below any storage rounding: the floor against which
the precision forms' deviations are read.

`run` advances one precision form, sampling Nu (the seven numbers of output_nu) and the kinetic
energy every --sample steps, and stores the samples, the device time and the final particle
state.  `compare` reports, for each form against the FP64 file: MLUPS, lattice HBM bandwidth at
the algorithmic bytes per node, Nu (mean over the second half of the run), KE, and the RMS
particle position and temperature differences.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

PRECISIONS = {"fp64": 0, "fp32": 1, "fp16s": 2, "bf16s": 3}
BYTES_PER_NODE = {0: 880.0, 1: 440.0, 2: 224.0, 3: 224.0}
HBM_PEAK_GBS = 8000.0


def load_rbtp():
    import importlib.util
    if "rbtp_amd" in sys.modules:
        return sys.modules["rbtp_amd"]
    pkg = os.path.join(ROOT, "rb-thermoparticles_amd")
    spec = importlib.util.spec_from_file_location("rbtp_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rbtp_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def workload(name, precision):
    from cases import config1, config2, config3, config5
    if name == "c1":
        return dict(config1(), precision=precision)
    return {"c2": config2, "c3": config3, "c5": config5}[name](precision=precision)


def run(args):
    from cases import local_particle_temperature
    rb = load_rbtp()
    prec = PRECISIONS[args.precision]
    kw = workload(args.workload, prec)
    if args.no_particles:
        kw.update(activate_particles=0, particle_fraction=0.0)
    case = rb.default_case(**kw)
    sim = rb.Sim(case)
    if args.shift_interval is not None:
        sim.set_shift_interval(args.shift_interval)
    T = rb.linear_temperature(case, 2024, 1e-3)
    if args.perturb:
        T = T + args.perturb * (rb.linear_temperature(case, 77, 1.0) -
                                rb.linear_temperature(case, 77, 0.0))
    sim.init(temperature=T)
    if sim.np:
        ppos = np.concatenate([sim.particles("pos_x"), sim.particles("pos_y"), sim.particles("pos_z")])
        sim.set_particles("temp", local_particle_temperature(T, ppos, case.lx, case.ly, case.lz))
    sim.synchronize()
    samples, ke, steps_at, traj = [], [], [], []
    dev_ms = 0.0
    wall0 = time.perf_counter()
    done = 0
    while done < args.steps:
        k = min(args.sample, args.steps - done)
        sim.step(k, macro_last=True)
        dev_ms += sim.timing().step_ms
        done += k
        nu = sim.nu()
        samples.append(nu)
        ke.append(sim.kinetic_energy())
        steps_at.append(done)
        if sim.np:
            traj.append(np.stack([sim.particles(k) for k in ("pos_x", "pos_y", "pos_z", "temp")]))
        print(f"[{args.workload} {args.precision}] step {done}: Nu_hot {nu[4]:.5f} Nu_bulk "
              f"{nu[2]:.5f} KE {ke[-1]:.4e} ({dev_ms / done:.3f} ms/step)", flush=True)
        if not np.all(np.isfinite(nu)):
            print("non-finite Nu: stopping", flush=True)
            break
    wall = time.perf_counter() - wall0
    parts = {name: sim.particles(name) for name in ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y",
                                                     "vel_z", "temp")}
    nlinks = len(sim.links()) if sim.np else 0
    np.savez(args.out, nu=np.array(samples), ke=np.array(ke), steps_at=np.array(steps_at),
             traj=np.array(traj) if traj else np.zeros((0, 4, 0)),
             dev_ms=dev_ms, wall_s=wall, n=sim.n, npart=sim.np, links=nlinks, precision=prec,
             workload=args.workload, perturb=args.perturb,
             shift_interval=-1 if args.shift_interval is None else args.shift_interval, **parts)
    print(f"[{args.workload} {args.precision}] {done} steps, {sim.n * done / (dev_ms * 1e-3) / 1e6:.0f} "
          f"MLUPS (device), {nlinks} links", flush=True)
    sim.close()


def compare(files):
    data = [dict(np.load(f)) for f in files]
    ref = data[0]
    if int(ref["precision"]) != 0:
        raise SystemExit("the first file must be the FP64 run")
    out = []
    for d in data:
        prec = int(d["precision"])
        steps = int(d["steps_at"][-1])
        mlups = float(d["n"]) * steps / (float(d["dev_ms"]) * 1e-3) / 1e6
        half = len(d["nu"]) // 2
        nu_mean = d["nu"][half:].mean(axis=0)
        nu_ref = ref["nu"][half:].mean(axis=0)
        row = dict(
            workload=str(d["workload"]), precision=[k for k, v in PRECISIONS.items() if v == prec][0],
            perturb=float(d.get("perturb", 0.0)), shift_interval=int(d.get("shift_interval", -1)),
            steps=steps, mlups=mlups,
            lattice_gbs_at_algorithmic_bytes=mlups * BYTES_PER_NODE[prec] / 1e3,
            mlups_vs_fp64=mlups / (float(ref["n"]) * int(ref["steps_at"][-1]) /
                                   (float(ref["dev_ms"]) * 1e-3) / 1e6),
            nu_hot_mean=float(nu_mean[4]), nu_cold_mean=float(nu_mean[3]), nu_bulk_mean=float(nu_mean[2]),
            nu_hot_rel_err=float(abs(nu_mean[4] - nu_ref[4]) / abs(nu_ref[4])),
            nu_cold_rel_err=float(abs(nu_mean[3] - nu_ref[3]) / abs(nu_ref[3])),
            nu_bulk_rel_err=float(abs(nu_mean[2] - nu_ref[2]) / max(abs(nu_ref[2]), 1e-300)),
            nu_hot_final=float(d["nu"][-1][4]),
            nu_hot_final_rel_err=float(abs(d["nu"][-1][4] - ref["nu"][-1][4]) / abs(ref["nu"][-1][4])),
            ke_final=float(d["ke"][-1]),
            ke_rel_err=float(abs(d["ke"][-1] - ref["ke"][-1]) / abs(ref["ke"][-1])),
            npart=int(d["npart"]), links=int(d["links"]))
        if int(d["npart"]) > 0:
            dpos = np.sqrt((d["pos_x"] - ref["pos_x"]) ** 2 + (d["pos_y"] - ref["pos_y"]) ** 2 +
                           (d["pos_z"] - ref["pos_z"]) ** 2)
            row.update(particle_rms_pos_err=float(np.sqrt(np.mean(dpos ** 2))),
                       particle_max_pos_err=float(dpos.max()),
                       particle_rms_temp_err=float(np.sqrt(np.mean((d["temp"] - ref["temp"]) ** 2))),
                       particle_temp_spread=float(ref["temp"].std()))
        if d.get("traj") is not None and ref.get("traj") is not None and d["traj"].size and ref["traj"].size:
            m = min(len(d["traj"]), len(ref["traj"]))
            dp = d["traj"][:m, :3] - ref["traj"][:m, :3]
            row["particle_rms_pos_err_by_sample"] = [float(x) for x in np.sqrt((dp ** 2).sum(axis=1).mean(axis=1))]
            row["particle_rms_temp_err_by_sample"] = [
                float(x) for x in np.sqrt(((d["traj"][:m, 3] - ref["traj"][:m, 3]) ** 2).mean(axis=1))]
            row["sample_steps"] = [int(x) for x in d["steps_at"][:m]]
        out.append(row)
        print(json.dumps(row), flush=True)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--workload", choices=["c1", "c2", "c3", "c5"], default="c5")
    r.add_argument("--precision", choices=list(PRECISIONS), default="fp64")
    r.add_argument("--steps", type=int, default=20000)
    r.add_argument("--sample", type=int, default=500)
    r.add_argument("--out", required=True)
    r.add_argument("--no-particles", action="store_true", help="the lattice alone")
    r.add_argument("--perturb", type=float, default=0.0,
                   help="amplitude of a second seeded noise added to the initial temperature")
    r.add_argument("--shift-interval", type=int, default=None,
                   help="steps between refreshes of the reduced-precision shift profile "
                        "(0: keep the conduction profile; default: the library's 500)")
    c = sub.add_parser("compare")
    c.add_argument("files", nargs="+")
    args = ap.parse_args()
    if args.cmd == "run":
        run(args)
    else:
        compare(args.files)


if __name__ == "__main__":
    main()
