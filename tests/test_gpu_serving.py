"""GPU tests of the serving batch (rbtp_batch) and the request loop.

Bar: a request's result is the same alone and in a mix.  Every node update, particle sum and
diagnostic reduction runs in a fixed order that does not depend on the batch, so the expected
agreement is bitwise; the assertions say so.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def case_kw(lx, ra, pr, npart, seed):
    from rbtp_amd.serve import request_case_kw
    kw = request_case_kw(lx, ra, pr, npart)
    kw["particle_rand_seed"] = seed
    return kw


class _Reqs:
    """Four requests of different size, Ra, Pr and particle count (built on first use)."""

    def __getitem__(self, i):
        if not hasattr(self, "_v"):
            self._v = [case_kw(32, 3e4, 0.71, 0, 1), case_kw(48, 2e5, 4.3, 6, 2),
                       case_kw(32, 1e6, 0.71, 3, 3), case_kw(64, 5e4, 0.71, 0, 4)]
        return self._v[i]


REQS = _Reqs()


def admit(rb, batch, kw, seed):
    case = rb.default_case(**kw)
    T = rb.linear_temperature(case, seed, 1e-3)
    return batch.admit(case, temperature=T)


def snapshot(sim):
    return {n: sim.populations(n) for n in ("f", "g")} | {"phi": sim.field("phi")}


@pytest.mark.parametrize("prec", [2, 0])
def test_alone_equals_mixed(rb, prec):
    alone = rb.Batch(precision=prec, max_slots=4)
    s = admit(rb, alone, REQS[1], 11)
    alone.step(60)
    ref = snapshot(alone.sim(s))
    nu_ref = alone.sim(s).nu()
    alone.close()
    mixed = rb.Batch(precision=prec, max_slots=8)
    others = [admit(rb, mixed, REQS[0], 10)]
    mixed.step(20)  # the mix is already running when the request arrives
    s2 = admit(rb, mixed, REQS[1], 11)
    others += [admit(rb, mixed, REQS[2], 12), admit(rb, mixed, REQS[3], 13)]
    mixed.step(40)
    mixed.release(others[0])
    mixed.step(20)
    got = snapshot(mixed.sim(s2))
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k
    assert np.array_equal(mixed.sim(s2).nu(), nu_ref)


def test_batch_matches_single_sim(rb):
    """A batched slot and a standalone rbtp_sim of the same case agree bitwise, also past the
    FP16 form's shift refreshes (every 500 steps): the batch refreshes each slot at the same
    steps, between graph replays, as rbtp_step does.  A standalone run without refreshes
    differs, so the comparison sees them."""
    kw = dict(REQS[2], precision=2)
    case = rb.default_case(**kw)
    T = rb.linear_temperature(case, 12, 1e-3)
    sim = rb.Sim(case)
    sim.init(temperature=T)
    sim.step(41)
    batch = rb.Batch(precision=2, max_slots=2)
    s = batch.admit(case, temperature=T)  # admission runs the first step
    batch.step(40)
    b = batch.sim(s)
    for k in ("f", "g"):
        assert np.array_equal(b.populations(k), sim.populations(k)), k
    for k in ("pos_x", "vel_x", "temp", "force_x"):
        assert np.array_equal(b.particles(k), sim.particles(k)), k
    # past the refreshes after steps 500 and 1000, in uneven chunks of graph replays and
    # single steps
    batch.step(437)
    batch.step(590)
    batch.step(33)
    sim.step(1060)
    plain = rb.Sim(case)
    plain.set_shift_interval(0)
    plain.init(temperature=T)
    plain.step(1101)
    for k in ("f", "g"):
        assert np.array_equal(b.populations(k), sim.populations(k)), f"{k} after refreshes"
    assert not np.array_equal(plain.populations("g"), sim.populations("g"))
    for k in ("pos_x", "vel_x", "temp", "force_x"):
        assert np.array_equal(b.particles(k), sim.particles(k)), k


@pytest.mark.parametrize("prec", [2, 0])
def test_ragged_slots_match_single_sims(rb, prec):
    """Slots whose node counts are not whole tiles (37 x 29 x 23 with four moving spheres,
    25 x 23 x 21 without), so that tiles of the batched launch end inside a slot: each slot
    equals a standalone rbtp_sim of its case bit for bit."""
    from cases import particle_case, single_phase
    cases = [
        (particle_case(37, 29, 23, npart=4, diam=6.0, ra=1e5, beta=0.5, precision=prec),
         np.asarray([(6.5, 7.0, 8.0), (30.2, 24.5, 18.6), (18.0, 14.5, 3.6),
                     (10.0, 22.0, 12.5)], dtype=np.float64).T.reshape(-1),
         np.array([0.3, -0.2, 0.0, 0.1])),
        (single_phase(25, 23, 21, ra=1e6, precision=prec), None, None),
    ]
    batch = rb.Batch(precision=prec, max_slots=4)
    sims, slots = [], []
    for i, (kw, ppos, ptemp) in enumerate(cases):
        case = rb.default_case(**kw)
        T = rb.linear_temperature(case, 30 + i, 1e-2)
        # arrays of the wrong length are refused before the library reads them
        with pytest.raises(ValueError):
            batch.admit(case, temperature=T[:-1], ppos=ppos, ptemp=ptemp)
        if ppos is not None:
            with pytest.raises(ValueError):
                batch.admit(case, temperature=T, ppos=ppos[:-1], ptemp=ptemp)
        slots.append(batch.admit(case, temperature=T, ppos=ppos, ptemp=ptemp))
        sim = rb.Sim(case)
        sim.init(temperature=T, ppos=ppos, ptemp=ptemp)
        sims.append(sim)
    batch.step(40)
    batch.step(9)
    for sim in sims:
        sim.step(50)  # the batch's admission ran each slot's first step
    for s, sim in zip(slots, sims):
        b = batch.sim(s)
        for k in ("f", "g"):
            assert np.array_equal(b.populations(k), sim.populations(k)), k
        assert np.array_equal(b.field("phi"), sim.field("phi"))
        if sim.np:
            for k in ("pos_x", "vel_z", "temp", "force_x", "heat"):
                assert np.array_equal(b.particles(k), sim.particles(k)), k
    batch.close()


def test_graph_replay_equals_eager(rb):
    res = []
    for graph_steps in (0, 10):
        batch = rb.Batch(precision=2, max_slots=4)
        slots = [admit(rb, batch, REQS[i], 20 + i) for i in range(3)]
        batch.set_launch(128, 512, graph_steps)
        batch.step(47)
        batch.step(33)
        res.append([snapshot(batch.sim(s)) for s in slots])
        batch.close()
    for a, b in zip(*res):
        for k in a:
            assert np.array_equal(a[k], b[k]), k


def test_tuning_leaves_state_unchanged(rb):
    res = []
    for tune in (False, True):
        batch = rb.Batch(precision=2, max_slots=4)
        slots = [admit(rb, batch, REQS[i], 30 + i) for i in range(3)]
        batch.step(10)
        if tune:
            blk, grd, ms = batch.tune()
            assert blk in (64, 128, 256) and grd >= 1 and ms > 0
        batch.step(30)
        res.append([snapshot(batch.sim(s)) for s in slots])
        batch.close()
    for a, b in zip(*res):
        for k in a:
            assert np.array_equal(a[k], b[k]), k


def test_admission_limits(rb):
    batch = rb.Batch(precision=2, max_slots=2)
    admit(rb, batch, REQS[0], 1)
    admit(rb, batch, REQS[0], 2)
    with pytest.raises(rb.CapacityError):
        admit(rb, batch, REQS[0], 3)
    tiny = rb.Batch(precision=2, max_slots=4, hbm_fraction=1e-6)
    with pytest.raises(rb.CapacityError):
        admit(rb, tiny, REQS[0], 4)


def test_request_loop_small(rb):
    from rbtp_amd.serve import make_requests, serve
    reqs = make_requests(6, mean_gap=0.05, seed=5, max_steps=600, max_particles=4,
                         sizes=(32, 48))
    for r in reqs:
        r.window = 200
    summary = serve(reqs, max_slots=4, chunk=100, tune=True)
    assert summary["requests"] == 6
    assert all(np.isfinite(r.latency) and r.latency > 0 for r in reqs)
    assert all(r.steps >= 201 for r in reqs)
    assert summary["p99_s"] >= summary["p50_s"] > 0
    assert summary["aggregate_mlups"] > 0


def test_full_size_requests_match_oracle(rb, O):
    """The largest requests the request generator makes (LX 96 and 128, 64 spheres each, the
    reference's rejection placement) served together in one FP64 batch, with a third small
    request beside them, agree with the oracle run of each case alone: populations bitwise,
    the particle state within 1e-12."""
    kws = [case_kw(128, 1e6, 0.71, 64, 5), case_kw(96, 3e5, 4.3, 64, 6)]
    batch = rb.Batch(precision=0, max_slots=4)
    slots = [admit(rb, batch, kw, 40 + i) for i, kw in enumerate(kws)]
    admit(rb, batch, case_kw(32, 3e4, 0.71, 0, 1), 39)
    batch.step(3)
    for kw, slot, seed in zip(kws, slots, (40, 41)):
        case = O.default_case(**kw)
        ora = O.Oracle(case)
        ora.init(temperature=O.linear_temperature(case, seed, 1e-3))
        ora.step(4)  # the admission's step and the three batched ones
        b = batch.sim(slot)
        assert b.np == 64
        for name in ("f", "g"):
            assert np.array_equal(b.populations(name), ora.populations(name)), (kw["lx"], name)
        for name in ("pos_x", "pos_y", "pos_z", "vel_x", "omega_z", "temp", "force_x", "heat"):
            a, r = b.particles(name), ora.particles(name)
            scale = max(float(np.max(np.abs(r))), 1e-300)
            assert float(np.max(np.abs(a - r))) / scale <= 1e-12, (kw["lx"], name)
        ora.close()
    batch.close()
