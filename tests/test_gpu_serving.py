"""Ensemble serving: requests of different sizes, mechanisms and step
counts batched into shared launches.  A request's result must not depend on
the mix it ran in (bitwise), the batch path must equal the level stepper on
the same problem (which is bitwise the oracle), and admission must respect
the memory budget."""
import ctypes as C

import numpy as np
import pytest

from particleclem_amd import serving as sv

pytestmark = pytest.mark.gpu


def small_stream(count, seed=11):
    return sv.request_stream(count, rate=0, seed=seed, n_lo=8, n_hi=20,
                             n_med=12, steps=(3, 9))


@pytest.mark.parametrize("mixed", [False, True])
def test_each_result_same_alone_and_in_the_mix(cuda, mixed):
    reqs = small_stream(8)
    srv = sv.Server(mixed=mixed, budget_bytes=4 << 30, keep_results=True)
    mix = {r.rid: r for r in srv.run(reqs, realtime=False)}
    srv.close()
    assert len(mix) == len(reqs)
    for r in reqs:
        one = sv.Request(r.rid, r.kind, r.n, r.nsteps, r.ratio, r.phi, 0.0, r.seed)
        s1 = sv.Server(mixed=mixed, budget_bytes=1 << 30, keep_results=True)
        got = s1.run([one], realtime=False)[0]
        s1.close()
        assert mix[r.rid].status == 0 and got.status == 0
        assert np.array_equal(got.result, mix[r.rid].result), (r.rid, r.kind, r.n)
    print(f"mixed={mixed}: {len(reqs)} requests bitwise equal alone and in the mix")


def test_batch_equals_level_stepper(cuda):
    """One tube and one slab request through the FP64 batch against
    pc_level_* on the same box, BCs and fields."""
    import torch
    from particleclem_amd.hydro import DeviceLevel, HydroContext, a4_from_tensor
    reqs = [sv.Request(0, sv.TUBE, 12, 5, 10.0, 0.8, 0.0, 5),
            sv.Request(1, sv.SLAB, 10, 4, 2.0, 1.1, 0.0, 6)]
    srv = sv.Server(mixed=False, budget_bytes=1 << 30, keep_results=True)
    out = {r.rid: r for r in srv.run(reqs, realtime=False)}
    srv.close()
    for r in reqs:
        P, bclo, bchi, mode = sv.request_fields(r)
        ctx = HydroContext(sv.NSPEC[r.kind])
        lev = DeviceLevel(ctx, [([0, 0, 0], [r.n - 1] * 3)], [0, 0, 0], [r.n - 1] * 3,
                          bclo, bchi, [1.0 / r.n] * 3, params=dict(ppm_type=1, cfl=0.3))
        lev._chk(lev.L.pc_level_prim_to_cons_v1(lev.h, 0, a4_from_tensor(P, [0, 0, 0]),
                                                 mode, None))
        lev.init_dt()
        lev.step(r.nsteps)
        ref = lev.read_box(0)
        nd = int(np.count_nonzero(ref[:, 4:-4, 4:-4, 4:-4] != out[r.rid].result[:, 4:-4, 4:-4, 4:-4]))
        print(f"request {r.rid} kind {r.kind}: {nd} values differ from the level stepper")
        assert nd == 0
        lev.close()
        ctx.close()


def test_mol_batch_equals_mol_level_and_mix_independence(cuda):
    """MOL requests (both Heun stages per step) through the FP64 batch:
    equal to the MOL level stepper on the same box, BCs and fields, and the
    same alone and in the mix."""
    from particleclem_amd.hydro import DeviceLevel, HydroContext, a4_from_tensor
    reqs = [sv.Request(0, sv.TUBE, 12, 5, 10.0, 0.8, 0.0, 5),
            sv.Request(1, sv.SLAB, 10, 4, 2.0, 1.1, 0.0, 6),
            sv.Request(2, sv.SLAB, 8, 3, 2.0, 0.6, 0.0, 7)]
    srv = sv.Server(mixed=False, budget_bytes=1 << 30, keep_results=True, mol=True)
    out = {r.rid: r for r in srv.run(reqs, realtime=False)}
    srv.close()
    for r in reqs:
        P, bclo, bchi, mode = sv.request_fields(r)
        ctx = HydroContext(sv.NSPEC[r.kind])
        lev = DeviceLevel(ctx, [([0, 0, 0], [r.n - 1] * 3)], [0, 0, 0], [r.n - 1] * 3,
                          bclo, bchi, [1.0 / r.n] * 3,
                          params=dict(ppm_type=1, cfl=0.3, do_mol=1))
        lev._chk(lev.L.pc_level_prim_to_cons_v1(lev.h, 0, a4_from_tensor(P, [0, 0, 0]),
                                                 mode, None))
        lev.init_dt()
        lev.step(r.nsteps)
        ref = lev.read_box(0)
        V = (slice(None), slice(4, -4), slice(4, -4), slice(4, -4))
        nd = int(np.count_nonzero(ref[V] != out[r.rid].result[V]))
        print(f"MOL request {r.rid} kind {r.kind}: {nd} values differ from the level stepper")
        assert out[r.rid].status == 0 and nd == 0
        lev.close()
        ctx.close()
        s1 = sv.Server(mixed=False, budget_bytes=1 << 30, keep_results=True, mol=True)
        one = sv.Request(r.rid, r.kind, r.n, r.nsteps, r.ratio, r.phi, 0.0, r.seed)
        got = s1.run([one], realtime=False)[0]
        s1.close()
        assert np.array_equal(got.result, out[r.rid].result)


@pytest.mark.parametrize("mol", [False, True], ids=["ctu", "mol"])
def test_mixed_batch_equals_mixed_level(cuda, mol):
    """The mixed-precision batch (FP32 intermediates and face fluxes) against
    the mixed-precision level stepper on the same box, BCs and fields: the
    same kernels and arithmetic, so bitwise equal, for the Godunov and the
    method-of-lines step."""
    from particleclem_amd.hydro import DeviceLevel, HydroContext, a4_from_tensor
    reqs = [sv.Request(0, sv.TUBE, 12, 5, 10.0, 0.8, 0.0, 5),
            sv.Request(1, sv.SLAB, 10, 4, 2.0, 1.1, 0.0, 6)]
    srv = sv.Server(mixed=True, budget_bytes=1 << 30, keep_results=True, mol=mol)
    out = {r.rid: r for r in srv.run(reqs, realtime=False)}
    srv.close()
    V = (slice(None), slice(4, -4), slice(4, -4), slice(4, -4))
    for r in reqs:
        P, bclo, bchi, mode = sv.request_fields(r)
        ctx = HydroContext(sv.NSPEC[r.kind])
        lev = DeviceLevel(ctx, [([0, 0, 0], [r.n - 1] * 3)], [0, 0, 0], [r.n - 1] * 3,
                          bclo, bchi, [1.0 / r.n] * 3,
                          params=dict(ppm_type=1, cfl=0.3, do_mol=int(mol)))
        lev.set_precision(True)
        lev._chk(lev.L.pc_level_prim_to_cons_v1(lev.h, 0, a4_from_tensor(P, [0, 0, 0]),
                                                 mode, None))
        lev.init_dt()
        lev.step(r.nsteps)
        ref = lev.read_box(0)
        nd = int(np.count_nonzero(ref[V] != out[r.rid].result[V]))
        print(f"mixed {'MOL' if mol else 'CTU'} request {r.rid} kind {r.kind}: {nd} values "
              f"differ from the mixed level stepper")
        assert out[r.rid].status == 0 and nd == 0
        lev.close()
        ctx.close()


def test_admission_under_a_tight_budget(cuda):
    reqs = small_stream(10, seed=3)
    probe = sv.Server(mixed=True, budget_bytes=64 << 20)
    per = max(probe.batch[r.kind].slot_bytes(r.n) for r in reqs)
    probe.close()
    budget = int(2.5 * per)
    srv = sv.Server(mixed=True, budget_bytes=budget)
    peak = [0]

    def tick(_, active):
        used = srv.batch[sv.SLAB].memory()[1]
        peak[0] = max(peak[0], used)
        assert used <= budget

    done = srv.run(reqs, realtime=False, on_tick=tick)
    srv.close()
    assert len(done) == len(reqs) and all(r.status == 0 for r in done)
    print(f"budget {budget >> 20} MiB, peak use {peak[0] >> 20} MiB, "
          f"{len(done)} requests served")


@pytest.mark.parametrize("max_slots", [512, 3])
def test_backfill_admission_matches_first_come_results(cuda, max_slots):
    """Backfilling under a tight budget (and, with 3 slots per batch, a
    request-slot limit that binds before memory does): every placement the
    admission plan predicts is where the arena puts the request (the server
    raises otherwise), the arena's free ranges and the live blocks tile it,
    no batch holds more than max_slots requests, and each request's result
    is bitwise the one it gets under first come first served (its
    arithmetic does not depend on when it was admitted)."""
    reqs = small_stream(16, seed=5)
    probe = sv.Server(mixed=True, budget_bytes=64 << 20)
    per = max(probe.batch[r.kind].slot_bytes(r.n) for r in reqs)
    probe.close()
    budget = int(2.5 * per)
    out, ticks = {}, {}
    for adm in ("fcfs", "backfill"):
        srv = sv.Server(mixed=True, budget_bytes=budget, keep_results=True,
                        admission=adm, bypass_depth=8, max_slots=max_slots)
        total = srv.batch[sv.SLAB].memory()[0]

        def tick(_, active, srv=srv, total=total):
            free = sum(b for _, b in srv.batch[sv.SLAB].free_ranges())
            held = sum(r.bytes for r in active)
            assert free + held == total
            for kind in srv.batch:
                assert sum(r.kind == kind for r in active) <= max_slots

        done = srv.run(small_stream(16, seed=5), realtime=False, on_tick=tick)
        ticks[adm] = srv.ticks
        srv.close()
        assert len(done) == len(reqs) and all(r.status == 0 for r in done)
        out[adm] = {r.rid: r for r in done}
    for rid, r in out["fcfs"].items():
        assert np.array_equal(out["backfill"][rid].result, r.result), rid
    order = lambda a: [r.rid for r in sorted(out[a].values(), key=lambda r: r.t_admit)]
    print(f"budget {budget >> 20} MiB, {max_slots} slots: ticks fcfs {ticks['fcfs']}, backfill "
          f"{ticks['backfill']}; admission order fcfs {order('fcfs')}, backfill "
          f"{order('backfill')}")


def test_ensemble_generator_shape_alone_equals_mix_under_budget(cuda):
    """Requests of the ensemble workload's own shape: 16 draws of the seeded
    generator (half gamma-law shock tubes with pressure ratio 2-20, half
    9-species premixed slabs with phi 0.5-1.2, sides lognormal in [32, 128]
    with median 64), mixed precision, step counts capped at 12 (the
    generator draws 10-200).  Served together under 80% of an arena that
    holds about five of the largest requests (so admission queues and
    backfills), no tick uses more than that budget, and every request's
    result is bitwise the one it gets served alone."""
    reqs = sv.request_stream(16, rate=0, seed=20261203)
    for r in reqs:
        r.nsteps = min(r.nsteps, 12)
    sides = sorted(r.n for r in reqs)
    assert 32 <= sides[0] and sides[-1] <= 128 and sides[0] < sides[-1]
    assert {r.kind for r in reqs} == {sv.TUBE, sv.SLAB}
    probe = sv.Server(mixed=True, budget_bytes=64 << 20)
    per = max(probe.batch[r.kind].slot_bytes(r.n) for r in reqs)
    probe.close()
    budget = int(0.8 * 5 * per)
    srv = sv.Server(mixed=True, budget_bytes=budget, keep_results=True)
    peak = [0]

    def tick(_, active):
        used = srv.batch[sv.SLAB].memory()[1]
        peak[0] = max(peak[0], used)
        assert used <= budget

    mix = {r.rid: r for r in srv.run(reqs, realtime=False, on_tick=tick)}
    ticks = srv.ticks
    srv.close()
    assert len(mix) == len(reqs) and all(r.status == 0 for r in mix.values())
    for r in reqs:
        one = sv.Request(r.rid, r.kind, r.n, r.nsteps, r.ratio, r.phi, 0.0, r.seed)
        s1 = sv.Server(mixed=True, budget_bytes=per + (64 << 20), keep_results=True)
        got = s1.run([one], realtime=False)[0]
        s1.close()
        assert got.status == 0
        assert np.array_equal(got.result, mix[r.rid].result), (r.rid, r.kind, r.n)
    print(f"16 generator requests (sides {sides[0]}-{sides[-1]}), budget "
          f"{budget >> 20} MiB, peak {peak[0] >> 20} MiB, {ticks} ticks: each "
          f"bitwise equal alone and in the mix")
