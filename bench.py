#!/usr/bin/env python3
"""Benchmark of the hot path: D3Q27 f+g thermal LBM time steps on one MI355X.

Workload: single-phase Rayleigh-Benard at the reference's production shape, 320x320x160 nodes,
Ra=1e8, Pr=0.71, FP64 (the plain drop-in form), synthetic initial state (linear temperature plus
seeded noise).  A "step" is one full time step of main.cu's sequence over all nodes.

Prints one JSON line.  value = lattice updates per second of the whole job (MLUPS), with the
state resident in HBM before the timed region starts.  Every form's lattice is created and
initialised before the probes run, so that each form's warm-up and timed steps follow device
work rather than the host's preparation (the device lowers its clocks when idle).  The line also
carries:
  roofline      the fused lattice kernel's algorithmic HBM bytes per launch (880 B/node: 54 FP64
                populations read and written plus F_x read and written) over its average launch
                time, measured with HIP events around every launch of the timed steps on the
                library's stream, beside two measured ceilings: a streaming copy and the same
                access pattern without arithmetic; kernel_ms_steady and frac_steady are the same
                kernel over 100 launches after the untimed Nu window, the device busy for seconds;
  cpu_baseline  the FP64 host oracle (OpenMP) timed on this machine's cores on a bounded sample,
                with the CPU model;
  forms         the FP16-storage (FP32 math), FP32 and BF16-storage forms on the same workload,
                with their Nu and KE deviations from FP64: at the end of the timed run and as the
                mean over an untimed window (10 000 steps by default) after it, checked against
                the Nu +-1 % bar; and the 300-sphere FP64 case, 5 000 timed steps by default;
  serving       the served batch (FP16 storage) under Poisson arrivals at half its measured
                capacity: time-to-result p50/p99 over 256 requests by default.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
BYTES_PER_NODE = {0: 880.0, 1: 440.0, 2: 224.0, 3: 224.0}
NU_BAR = 0.01  # a reduced-precision form is usable when its Nu stays within 1 % of FP64


def cpu_model():
    """The host CPU's model name (first processor in /proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def load_rbtp():
    import importlib.util
    if "rbtp_amd" in sys.modules:
        return sys.modules["rbtp_amd"]
    spec = importlib.util.spec_from_file_location(
        "rbtp_amd", os.path.join(ROOT, "rb-thermoparticles_amd", "__init__.py"),
        submodule_search_locations=[os.path.dirname(os.path.join(ROOT, "rb-thermoparticles_amd", "__init__.py"))])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rbtp_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def measured_traffic(name):
    """HBM bytes per launch and their ratio to the algorithmic bytes, from the committed PMC
    summary profiles/lattice_<name>_320x320x160_traffic.json, or (None, None)."""
    path = os.path.join(ROOT, "profiles", f"lattice_{name}_320x320x160_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        t = json.load(fh)
    return t.get("bytes_per_launch"), t.get("traffic_over_algorithmic")


def prepare_form(rb, kw):
    """A lattice of one precision form, created and initialised (host work: the device idles)."""
    case = rb.default_case(**kw)
    sim = rb.Sim(case)
    sim.init(temperature=rb.linear_temperature(case, 2024, 1e-3))
    return sim


def run_form(sim, steps, warmup, log, nu_window=0):
    n = sim.n
    sim.step(warmup, macro_last=False)
    sim.synchronize()
    # HIP events around every lattice launch of the timed steps, on the library's stream; they
    # are recorded without waiting and read after the run's synchronisation
    sim.set_kernel_timing(True)
    t0 = time.perf_counter()
    sim.step(steps, macro_last=True)
    t1 = time.perf_counter()
    wall_ms = (t1 - t0) * 1e3
    tm = sim.timing()
    sim.set_kernel_timing(False)
    dev_ms = tm.step_ms
    nu = sim.nu()
    ke = sim.kinetic_energy()
    # untimed continuation: Nu sampled every 100 steps for a window average
    nu_samples = []
    for _ in range(nu_window // 100):
        sim.step(100, macro_last=True)
        nu_samples.append(sim.nu())
    nu_avg = np.mean(nu_samples, axis=0) if nu_samples else nu
    # the kernel again after the window, the device busy for seconds: its steady-state time
    steady = None
    if nu_window:
        sim.set_kernel_timing(True)
        sim.step(100, macro_last=False)
        steady = sim.timing().kernel_ms
        sim.set_kernel_timing(False)
    prec = sim.case.precision
    log(f"form precision={prec}: {steps} steps in {wall_ms:.1f} ms wall "
        f"({dev_ms:.1f} ms device), kernel avg {tm.kernel_ms:.4f} ms over {tm.launches}")
    return dict(n=n, wall_ms=wall_ms, dev_ms=dev_ms, kernel_ms=tm.kernel_ms, kernel_ms_steady=steady,
                kernel_launches=tm.launches, nu=nu.tolist(), ke=ke, nu_avg=np.asarray(nu_avg).tolist(),
                nu_window=nu_window)


def cpu_baseline(kw, log, budget_s=20.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    case = O.default_case(**kw)
    ora = O.Oracle(case)
    T = O.linear_temperature(case, 2024, 1e-3)
    ora.init(temperature=T)
    ora.step(1)  # warm-up (page faults, thread start)
    t0 = time.perf_counter()
    steps = 0
    while True:
        ora.step(1)
        steps += 1
        if time.perf_counter() - t0 > budget_s or steps >= 50:
            break
    dt = time.perf_counter() - t0
    n = ora.n
    ora.close()
    mlups = n * steps / dt / 1e6
    log(f"cpu oracle: {steps} steps of {n} nodes in {dt:.2f} s -> {mlups:.2f} MLUPS "
        f"on {O.threads()} threads")
    return dict(value=mlups, unit="MLUPS", cores=O.threads(), kind="port",
                cpu_model=cpu_model(), logical_cpus=os.cpu_count(),
                sample=f"FP64 oracle, {steps} steps (after 1 warm-up) of the same "
                       f"{kw['lx']}x{kw['ly']}x{kw['lz']} Ra=1e8 workload")


def run_particles(rb, steps, warmup, log):
    """Particle-laden production case (320x320x160, 300 spheres): MLUPS of the full step."""
    from cases import config3, local_particle_temperature
    kw = config3(precision=0)
    case = rb.default_case(**kw)
    sim = rb.Sim(case)
    T = rb.linear_temperature(case, 2024, 1e-3)
    sim.init(temperature=T)  # reference rejection placement (glibc rand, seed 86143)
    ppos = np.concatenate([sim.particles("pos_x"), sim.particles("pos_y"), sim.particles("pos_z")])
    sim.set_particles("temp", local_particle_temperature(T, ppos, case.lx, case.ly, case.lz))
    sim.step(warmup, macro_last=False)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(steps, macro_last=True)
    dt = time.perf_counter() - t0
    n = sim.n
    nlinks = len(sim.links())
    nu = sim.nu()
    sim.close()
    mlups = n * steps / dt / 1e6
    log(f"particle-laden: {sim.np} spheres, {nlinks} links, {steps} steps in {dt*1e3:.1f} ms "
        f"-> {mlups:.0f} MLUPS")
    return dict(mlups=mlups, ms_per_step=dt * 1e3 / steps, npart=int(sim.np), links=nlinks,
                steps=steps, nu=nu.tolist(), workload="320x320x160 Ra=1e8 + 300 spheres d=10, FP64")


def run_serving(rb, nreq, log):
    """Served batch: Poisson arrivals at 50% of the measured batch capacity."""
    from rbtp_amd.serve import make_requests, serve
    # capacity: aggregate rate of a full batch of typical requests, then arrivals at half of it
    probe = make_requests(16, mean_gap=0.0, seed=99, max_steps=1000)
    cal = serve(probe, max_slots=64, chunk=100)
    mean_updates = np.mean([r.nodes for r in probe]) * 5000
    capacity = cal["aggregate_mlups"] * 1e6 / mean_updates  # requests per second
    mean_gap = 1.0 / (0.5 * capacity)
    log(f"serving calibration: {cal['aggregate_mlups']:.0f} MLUPS, capacity {capacity:.2f} req/s, "
        f"mean gap {mean_gap:.3f} s")
    reqs = make_requests(nreq, mean_gap=mean_gap, seed=2024)
    out = serve(reqs, max_slots=64, chunk=100, log=log)
    out.update(mean_gap_s=mean_gap, load=0.5, precision="fp16s",
               workload=f"{nreq} requests, Poisson arrivals (seed 2024) at half the calibrated "
                        f"capacity; cavities LX in {{32,48,64,96,128}}, LY=LX, LZ=LX/2, Ra "
                        f"log-uniform in [1e4,1e7], Pr in {{0.71,4.3}}, 0-64 spheres; each "
                        f"retired at Nu change < 1e-4 per 500 steps or 5000 steps")
    log(f"serving: p50 {out['p50_s']:.2f} s, p99 {out['p99_s']:.2f} s, "
        f"{out['aggregate_mlups']:.0f} MLUPS over {out['wall_s']:.1f} s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--nu-window", type=int, default=10000,
                    help="untimed steps after the timed run over which Nu is averaged, for "
                         "FP64 and every reduced-precision form")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the FP16-, FP32- and BF16-storage forms")
    ap.add_argument("--no-particles", action="store_true", help="skip the particle-laden case")
    ap.add_argument("--particle-steps", type=int, default=5000,
                    help="timed steps of the 300-sphere case (after a tenth as many warm-up "
                         "steps), independent of --steps")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--batch-probe", type=int, default=0, metavar="K",
                    help="K eager (not graph-replayed) steps of a 16-request serving batch, "
                         "for kernel traces; no JSON")
    ap.add_argument("--counter-run", action="store_true",
                    help="minimal run for memory-counter passes: 3 lattice steps at "
                         "320x320x160 and one 1 GiB copy-probe launch, no JSON")
    ap.add_argument("--counter-precision", type=int, default=0, choices=(0, 1, 2, 3),
                    help="precision form of the counter run's lattice steps (0 FP64, 1 FP32, "
                         "2 FP16 storage, 3 BF16 storage)")
    ap.add_argument("--serve", type=int, default=256,
                    help="requests of the served-batch workload (0: skip it)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is a design, not code)", file=sys.stderr)
        sys.exit(2)

    def log(msg):
        if not args.quiet:
            print(f"[bench] {msg}", file=sys.stderr, flush=True)

    from cases import config2
    rb = load_rbtp()
    kw64 = config2(precision=0)
    if args.batch_probe:
        # eager steps of a 16-request serving batch (FP16 storage, 0-64 spheres each)
        from rbtp_amd.serve import admit_request, make_requests
        batch = rb.Batch(precision=rb.PREC_FP16S, max_slots=64)
        reqs = make_requests(16, mean_gap=0.0, seed=99, max_steps=1000)
        for r in reqs:
            admit_request(batch, r)
        batch.set_launch(256, 16384, 0)
        batch.step(10)
        batch.step(args.batch_probe)
        ms = batch.info()["last_ms"]
        log(f"batch probe: {len(reqs)} requests, {sum(r.nodes for r in reqs)} nodes, "
            f"{ms / args.batch_probe:.3f} ms per batched step")
        batch.close()
        return
    if args.counter_run:
        case = rb.default_case(**config2(precision=args.counter_precision))
        sim = rb.Sim(case)
        sim.init(temperature=rb.linear_temperature(case, 2024, 1e-3))
        sim.step(3, macro_last=False)
        sim.close()
        rb.bandwidth_probe(1 << 30, 1)
        log("counter run done")
        return
    # Every lattice is set up first, and freed only after the last form has run: the host-side
    # preparation and the freeing of a lattice, during which the device idles and lowers its
    # clocks, then lie outside the forms' runs.  The forms run back to back, FP64 last, so that
    # each form's warm-up and timed steps follow the previous form's steps (the first, the
    # probes).
    form_precs = () if args.no_forms else (("fp16s", 2), ("fp32", 1), ("bf16s", 3))
    sims = {0: prepare_form(rb, kw64)}
    for _, prec in form_precs:
        sims[prec] = prepare_form(rb, config2(precision=prec))
    copy_gbs = rb.bandwidth_probe(2 << 30, 20)
    log(f"copy probe: {copy_gbs:.0f} GB/s")
    stencil_gbs = rb.stencil_probe(kw64["lx"], kw64["ly"], kw64["lz"], 20)
    log(f"stencil-pattern probe: {stencil_gbs:.0f} GB/s")
    # the alternative streaming scheme (in-place AA) as access patterns on the same lattice
    aa_even = rb.streaming_probe(1, kw64["lx"], kw64["ly"], kw64["lz"], 20)
    aa_odd = rb.streaming_probe(2, kw64["lx"], kw64["ly"], kw64["lz"], 20)
    aa_pair = 2.0 / (1.0 / aa_even + 1.0 / aa_odd)  # mean rate over an even + odd step pair
    log(f"AA-pattern probes: even {aa_even:.0f}, odd {aa_odd:.0f}, pair {aa_pair:.0f} GB/s")
    # the other candidate pull patterns: FP64 with plain 8-byte loads, and the 16-bit storage
    # forms' plain pull (their kernels' pattern), row pull with lane shifts, two nodes per thread
    probes = {name: rb.streaming_probe(scheme, kw64["lx"], kw64["ly"], kw64["lz"], 20)
              for name, scheme in (("fp64_plain_pull", 3), ("p16_plain_pull", 4),
                                   ("p16_row_pull", 5), ("p16_pair_pull", 6))}
    log("pattern probes: " + ", ".join(f"{k} {v:.0f}" for k, v in probes.items()) + " GB/s")
    runs = {}
    for prec in (1, 2, 3, 0):
        if prec in sims:
            runs[prec] = run_form(sims[prec], args.steps, args.warmup, log, args.nu_window)
    for sim in sims.values():
        sim.close()
    r64 = runs[0]
    n = r64["n"]
    mlups = n * args.steps / (r64["wall_ms"] * 1e-3) / 1e6
    achieved = BYTES_PER_NODE[0] * n / (r64["kernel_ms"] * 1e-3) / 1e9
    forms = {}
    if form_precs:
        for name, prec in form_precs:
            window = args.nu_window
            r = runs[prec]
            m = n * args.steps / (r["wall_ms"] * 1e-3) / 1e6
            form = dict(
                mlups=m, ratio_vs_fp64=m / mlups, kernel_ms=r["kernel_ms"],
                kernel_gbs=BYTES_PER_NODE[prec] * n / (r["kernel_ms"] * 1e-3) / 1e9,
                nu_bulk=r["nu"][2], nu_hot=r["nu"][4],
                nu_hot_rel_err_vs_fp64=abs(r["nu"][4] - r64["nu"][4]) / abs(r64["nu"][4]),
                ke_rel_err_vs_fp64=abs(r["ke"] - r64["ke"]) / abs(r64["ke"]),
                bytes_per_node=BYTES_PER_NODE[prec])
            if r["kernel_ms_steady"]:
                form.update(kernel_ms_steady=r["kernel_ms_steady"],
                            kernel_gbs_steady=BYTES_PER_NODE[prec] * n /
                            (r["kernel_ms_steady"] * 1e-3) / 1e9)
            if prec == 2:
                # HBM bytes per launch from the committed PMC passes of this form
                form["traffic"], form["traffic_over_algorithmic"] = measured_traffic("fp16s")
            if prec in (2, 3):
                # the 16-bit kernels pull with plain 2-byte loads: their pattern's ceiling
                form["pattern_probe_gbs"] = probes["p16_plain_pull"]
                form["frac_of_pattern_probe"] = (form.get("kernel_gbs_steady", form["kernel_gbs"])
                                                 / probes["p16_plain_pull"])
            if window:
                err = abs(r["nu_avg"][4] - r64["nu_avg"][4]) / abs(r64["nu_avg"][4])
                form.update(
                    nu_hot_window_avg=r["nu_avg"][4], nu_hot_window_avg_fp64=r64["nu_avg"][4],
                    nu_hot_window_rel_err_vs_fp64=err, nu_window_steps=window)
            else:
                err = form["nu_hot_rel_err_vs_fp64"]
            # the Nu bar, on the window mean where there is one
            form["nu_bar"] = {"tolerance": NU_BAR, "rel_err": err,
                              "meets": bool(err <= NU_BAR and
                                            form["nu_hot_rel_err_vs_fp64"] <= NU_BAR)}
            log(f"form {name}: Nu_hot error vs FP64 {err:.2e} "
                f"({'within' if form['nu_bar']['meets'] else 'FAILS'} the {NU_BAR:.0%} bar)")
            forms[name] = form
    if not args.no_particles:
        forms["fp64_particle_laden"] = run_particles(rb, max(args.particle_steps, 1),
                                                     max(args.particle_steps // 10, 1), log)
    traffic, _ = measured_traffic("fp64")
    cpu = None if args.no_cpu else cpu_baseline(kw64, log)
    serving = run_serving(rb, args.serve, log) if args.serve > 0 else None
    out = {
        "metric": "MLUPS (D3Q27 f+g), FP64, 320x320x160 Ra=1e8",
        "value": mlups,
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["wall_ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (linear temperature + 1e-3 seeded noise, fluid at rest)",
        "config": {"workload": "rb_single_phase_320x320x160_ra1e8_pr0.71_fp64",
                   "lx": 320, "ly": 320, "lz": 160, "rayleigh": 1e8, "prandtl": 0.71,
                   "visc": kw64["visc"], "precision": "fp64", "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": ("k_lattice_step_rows" if kw64["lx"] % 64 == 0 and
                                (n % 256) == 0 else "k_lattice_step<double,double,strict>"),
                     "copy_probe_gbs": copy_gbs, "frac_of_copy_probe": achieved / copy_gbs,
                     "stencil_probe_gbs": stencil_gbs,
                     "frac_of_stencil_probe": achieved / stencil_gbs,
                     "aa_pattern_probe_gbs": {"even": aa_even, "odd": aa_odd, "pair": aa_pair},
                     "pattern_probes_gbs": probes,
                     "kernel_ms": r64["kernel_ms"], "bytes_per_node": BYTES_PER_NODE[0],
                     # the same kernel after the untimed Nu window (device busy for seconds);
                     # the timed steps follow the host's preparation of the initial state, and
                     # the device's clocks take tens of milliseconds of load to come back up
                     "kernel_ms_steady": r64["kernel_ms_steady"],
                     "frac_steady": (BYTES_PER_NODE[0] * n / (r64["kernel_ms_steady"] * 1e-3)
                                     / 1e9 / HBM_PEAK_GBS) if r64["kernel_ms_steady"] else None},
        "cpu_baseline": cpu,
        "forms": forms,
        "serving": serving,
        "nu_fp64": r64["nu"],
        "ke_fp64": r64["ke"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
