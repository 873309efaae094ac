"""Continuous-batching request loop over the serving batch (rbtp_batch).

Requests (cavity solves of uneven size, Ra, Pr and particle count) arrive over time.  The loop
admits every arrived request as soon as a slot and HBM are free, advances all resident requests
together in chunks of batched steps (hipGraph replays inside rbtp_batch_step), checks each
request's Nu every `window` steps, and retires a request when Nu changed by less than `tol` over
the last window or it reached `max_steps`.  The Nu watched is the mean of the two wall Nusselt
numbers of output_nu: the bulk value also counts the velocity that refill_new_fluid_kernel
leaves at refilled nodes (momentum over the density deviation), which keeps it noisy while
spheres move.  Time-to-result is measured from arrival to retirement.  Per-request results do not
depend on what else is in the batch: every node update, particle sum and Nu reduction runs in a
fixed order independent of the batch composition.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import Batch, CapacityError, PREC_FP16S, default_case, linear_temperature

SIZES = (32, 48, 64, 96, 128)


@dataclass
class Request:
    rid: int
    case_kw: dict
    arrival: float
    seed: int = 0
    max_steps: int = 5000
    window: int = 500
    tol: float = 1e-4
    admitted: float = math.nan
    finished: float = math.nan
    steps: int = 0
    nodes: int = 0
    nu_history: list = field(default_factory=list)
    result: dict = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def request_case_kw(lx: int, ra: float, pr: float, npart: int) -> dict:
    """A served cavity: LX=LY, LZ=LX/2, T_hot=+0.5, T_cold=-0.5, free-fall velocity 0.1414
    (VISC = 0.1414*LX*sqrt(Pr/Ra)), spheres of diameter LX/12 with the reference's density and
    heat-capacity ratio 1 (neutrally buoyant)."""
    kw = dict(lx=lx, ly=lx, lz=lx // 2, rayleigh=ra, prandtl=pr, t_hot=0.5, t_cold=-0.5,
              t_ref=0.0, visc=0.1414 * lx * math.sqrt(pr / ra), beta=1.0,
              activate_particles=1 if npart > 0 else 0, npart_override=npart,
              particle_diam_frac=1.0 / 12.0, particle_fraction=0.0, rho_particle_ratio=1.0,
              cp_particle_ratio=1.0)
    return kw


def make_requests(n: int, mean_gap: float, seed: int = 0, max_steps: int = 5000,
                  max_particles: int = 64, sizes=SIZES) -> list:
    """Poisson arrivals (exponential gaps of mean `mean_gap` s); LX uniform over `sizes`,
    Ra log-uniform in [1e4, 1e7], Pr in {0.71, 4.3}, 0..max_particles spheres."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(mean_gap) if mean_gap > 0 else 0.0
        lx = int(rng.choice(sizes))
        ra = float(10.0 ** rng.uniform(4.0, 7.0))
        pr = float(rng.choice([0.71, 4.3]))
        npart = int(rng.integers(0, max_particles + 1))
        out.append(Request(rid=i, case_kw=request_case_kw(lx, ra, pr, npart), arrival=t,
                           seed=int(rng.integers(1, 2 ** 31)), max_steps=max_steps))
    return out


def admit_request(batch: Batch, r: Request) -> int:
    """Admits request r into the batch (seeded linear temperature); returns its slot."""
    kw = dict(r.case_kw)
    kw["particle_rand_seed"] = r.seed % 2147483647
    case = default_case(**kw)
    T = linear_temperature(case, r.seed, 1e-3)
    slot = batch.admit(case, temperature=T)
    r.nodes = case.lx * case.ly * case.lz
    return slot


def serve(requests, precision: int = PREC_FP16S, max_slots: int = 64, chunk: int = 100,
          hbm_fraction: float = 0.8, tune: bool = True, log=None, clock=time.perf_counter):
    """Run the request loop to completion; returns a summary dict and fills each Request."""
    if any(r.window % chunk for r in requests):
        raise ValueError("chunk must divide every request's window")
    batch = Batch(precision=precision, max_slots=max_slots, hbm_fraction=hbm_fraction)
    pending = deque(sorted(requests, key=lambda r: r.arrival))
    active = {}
    updates = 0
    tuned_at = None
    tunes = 0
    t0 = clock()
    while pending or active:
        now = clock() - t0
        changed = False
        while pending and pending[0].arrival <= now:
            r = pending[0]
            try:
                slot = admit_request(batch, r)
            except CapacityError:
                break
            pending.popleft()
            r.admitted = clock() - t0
            r.steps = 1
            updates += r.nodes
            active[slot] = r
            changed = True
        if not active:
            if pending:
                time.sleep(max(0.0, min(0.05, pending[0].arrival - (clock() - t0))))
            continue
        if tune and changed and (tuned_at is None or clock() - tuned_at > 2.0):
            blk, grd, ms = batch.tune()
            tuned_at = clock()
            tunes += 1
            if log:
                log(f"tuned on {len(active)} slots: tile {blk}, grid {grd}, {ms:.3f} ms/step")
        batch.step(chunk, macro_last=True)
        done = []
        for slot, r in active.items():
            r.steps += chunk
            updates += r.nodes * chunk
            if (r.steps - 1) % r.window == 0:
                sim = batch.sim(slot)
                nu = sim.nu()
                r.nu_history.append(0.5 * float(nu[3] + nu[4]))
                converged = (len(r.nu_history) >= 2 and
                             abs(r.nu_history[-1] - r.nu_history[-2]) < r.tol)
                if converged or r.steps >= r.max_steps:
                    r.result = dict(nu=nu.tolist(), ke=sim.kinetic_energy(), steps=r.steps,
                                    converged=bool(converged))
                    r.finished = clock() - t0
                    done.append(slot)
        for slot in done:
            batch.release(slot)
            del active[slot]
    wall = clock() - t0
    batch.close()
    lat = np.array([r.latency for r in requests])
    return dict(requests=len(requests), wall_s=wall, aggregate_mlups=updates / wall / 1e6,
                updates=updates, p50_s=float(np.percentile(lat, 50)),
                p99_s=float(np.percentile(lat, 99)), mean_s=float(lat.mean()),
                converged=sum(1 for r in requests if r.result.get("converged")), tunes=tunes)
