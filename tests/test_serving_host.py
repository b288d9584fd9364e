"""Host logic of the serving loop (CPU): the seeded request stream and the
latency summary."""
import numpy as np
import pytest

from particleclem_amd import serving as sv


def test_request_stream_is_seeded_and_within_bounds():
    a = sv.request_stream(500, rate=100.0, seed=5)
    b = sv.request_stream(500, rate=100.0, seed=5)
    c = sv.request_stream(500, rate=100.0, seed=6)
    key = lambda rs: [(r.kind, r.n, r.nsteps, r.ratio, r.phi, r.t_arrive, r.seed) for r in rs]
    assert key(a) == key(b) and key(a) != key(c)
    n = np.array([r.n for r in a])
    assert n.min() >= 32 and n.max() <= 128 and 56 <= np.median(n) <= 72
    st = np.array([r.nsteps for r in a])
    assert st.min() >= 10 and st.max() <= 200
    assert all(2.0 <= r.ratio <= 20.0 and 0.5 <= r.phi <= 1.2 for r in a)
    kinds = np.array([r.kind for r in a])
    assert 0.4 < kinds.mean() < 0.6
    t = np.array([r.t_arrive for r in a])
    assert np.all(np.diff(t) >= 0)
    # Poisson at 100/s: mean gap 10 ms
    assert 0.008 < np.diff(t).mean() < 0.012
    assert len({r.seed for r in a}) == len(a)


def test_summary_percentiles():
    reqs = sv.request_stream(100, rate=50.0, seed=1)
    for i, r in enumerate(reqs):
        r.t_admit = r.t_arrive + 0.001
        r.t_done = r.t_arrive + 0.01 * (i + 1)
    s = sv.summarize(reqs, wall=10.0)
    assert abs(s["latency_ms"]["p50"] - 505.0) < 1e-6
    assert abs(s["latency_ms"]["p99"] - 990.1) < 1e-6
    assert s["requests"] == 100
    cells = sum(r.n ** 3 * r.nsteps for r in reqs)
    assert abs(s["mcell_updates_per_s"] - cells / 10.0 / 1e6) < 1e-9


class _Units:
    """Brute-force arena: a map of 256-byte units; a request takes the
    lowest run of free units long enough for it."""

    def __init__(self, nunits):
        self.free = np.ones(nunits, dtype=bool)

    def _runs(self):
        d = np.diff(np.concatenate(([0], self.free.astype(np.int8), [0])))
        return np.flatnonzero(d == 1), np.flatnonzero(d == -1)

    def alloc(self, nbytes):
        need = -(-nbytes // 256)
        s, e = self._runs()
        ok = np.flatnonzero(e - s >= need)
        if not len(ok):
            return -1
        u = int(s[ok[0]])
        assert self.free[u:u + need].all()
        self.free[u:u + need] = False
        return u * 256

    def release(self, off, nbytes):
        u, need = off // 256, -(-nbytes // 256)
        assert not self.free[u:u + need].any()
        self.free[u:u + need] = True

    def fits(self, nbytes):
        s, e = self._runs()
        return bool(np.any(e - s >= -(-nbytes // 256)))

    def ranges(self):
        s, e = self._runs()
        return [(int(a) * 256, int(b - a) * 256) for a, b in zip(s, e)]


def test_arena_model_places_as_first_fit_over_units():
    rng = np.random.default_rng(11)
    ref = _Units(4096)
    m = sv.ArenaModel([(0, 4096 * 256)])
    live = []
    for _ in range(3000):
        if live and rng.random() < 0.45:
            off, b = live.pop(int(rng.integers(len(live))))
            ref.release(off, b)
            m.release(off, b)
        else:
            b = int(rng.integers(1, 300 * 256))
            assert m.fits(b) == ref.fits(b)
            o = ref.alloc(b)
            assert m.alloc(b) == o
            if o >= 0:
                live.append((o, b))
        assert m.fr == ref.ranges()


def test_ticks_until_fit_against_stepping_the_clock():
    rng = np.random.default_rng(12)
    for trial in range(400):
        ref = _Units(1024)
        held = []
        for _ in range(int(rng.integers(1, 40))):
            b = int(rng.integers(1, 80 * 256))
            o = ref.alloc(b)
            if o >= 0:
                held.append((int(rng.integers(1, 30)), o, b, int(rng.integers(2))))
        m = sv.ArenaModel(ref.ranges())
        need = int(rng.integers(1, 1100 * 256))
        kind = int(rng.integers(2))
        slots = None if trial % 2 else {0: int(rng.integers(-3, 2)), 1: int(rng.integers(-3, 2))}
        got = sv.ticks_until_fit(m, held, need, kind, slots)
        want, free = None, (slots[kind] if slots else 1)
        for t in range(0, 31):
            for left, o, b, k in held:
                if left == t:
                    ref.release(o, b)
                    free += k == kind
            if free > 0 and ref.fits(need):
                want = t
                break
        assert got == want


class _Req:
    def __init__(self, rid, nbytes, nsteps, arrive, kind):
        self.rid, self.bytes, self.nsteps, self.arrive = rid, nbytes, nsteps, arrive
        self.kind = kind
        self.done = 0
        self.off = self.admit = None


def _simulate(reqs, arena, depth, max_slots):
    """Tick loop of the serving queue on the brute-force arena with
    max_slots request slots per kind; admissions from backfill_plan (depth
    0: first come first served).  Returns the admission ticks and, per head
    that had to wait, (first tick blocked, reserved tick)."""
    ref = _Units(arena // 256)
    pending = sorted(reqs, key=lambda r: r.arrive)
    waiting, active, promised = [], [], {}
    tick = 0
    while pending or waiting or active:
        while pending and pending[0].arrive <= tick:
            waiting.append(pending.pop(0))
        held = [(r.nsteps - r.done, r.off, r.bytes, r.kind) for r in active]
        slots = {q: max_slots - sum(r.kind == q for r in active) for q in (0, 1)}
        assert min(slots.values()) >= 0
        # the head left after the first-come admissions, and its reservation
        probe, ph, pf = sv.ArenaModel(ref.ranges()), list(held), dict(slots)
        k = 0
        while (k < len(waiting) and pf[waiting[k].kind] > 0
               and probe.fits(waiting[k].bytes)):
            o = probe.alloc(waiting[k].bytes)
            ph.append((waiting[k].nsteps, o, waiting[k].bytes, waiting[k].kind))
            pf[waiting[k].kind] -= 1
            k += 1
        if k < len(waiting) and waiting[k].rid not in promised:
            s = sv.ticks_until_fit(probe, ph, waiting[k].bytes, waiting[k].kind, pf)
            assert s is not None
            promised[waiting[k].rid] = (tick, tick + s)
        plan = sv.backfill_plan(waiting, held, sv.ArenaModel(ref.ranges()), depth, slots)
        for j, off in plan:
            r = waiting[j]
            r.off = ref.alloc(r.bytes)
            assert r.off == off
            r.admit = tick
            active.append(r)
        for j, _ in sorted(plan, reverse=True):
            del waiting[j]
        for r in active:
            r.done += 1
        for r in [r for r in active if r.done == r.nsteps]:
            ref.release(r.off, r.bytes)
        active = [r for r in active if r.done < r.nsteps]
        tick += 1
    return {r.rid: r.admit for r in reqs}, promised


def _stream(seed, count=300):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(0.5, size=count)).astype(int)
    sides = np.clip(np.round(64 * np.exp(0.5 * rng.standard_normal(count))), 32, 128)
    return [_Req(i, int((sides[i] / 16) ** 3 * 256) + int(rng.integers(0, 256)),
                 int(rng.integers(10, 201)), int(t[i]), int(rng.integers(2)))
            for i in range(count)]


def test_backfill_never_delays_the_head_and_cuts_waits():
    arena = 8000 * 256   # about 16 of the largest requests
    # memory binds alone (1000 slots), slots bind alone (9), both (24, 28)
    for seed, max_slots in ((1, 1000), (2, 1000), (5, 9), (3, 24), (6, 28)):
        fc, _ = _simulate(_stream(seed), arena, 0, max_slots)
        reqs = _stream(seed)
        bf, promised = _simulate(reqs, arena, 64, max_slots)
        assert all(v is not None for v in bf.values())
        assert promised
        for rid, (blocked, reserved) in promised.items():
            assert blocked <= bf[rid] <= reserved, (seed, rid, blocked, bf[rid], reserved)
        arrive = {r.rid: r.arrive for r in reqs}
        wf = np.array([fc[i] - arrive[i] for i in arrive])
        wb = np.array([bf[i] - arrive[i] for i in arrive])
        print(f"seed {seed}, {max_slots} slots per kind: wait ticks fcfs p50 {np.median(wf):.0f} "
              f"p99 {np.percentile(wf, 99):.0f} mean {wf.mean():.1f}; backfill "
              f"p50 {np.median(wb):.0f} p99 {np.percentile(wb, 99):.0f} "
              f"mean {wb.mean():.1f}")
        assert wb.mean() <= wf.mean()


class _FakeBatch:
    """Host stand-in for a Batch: an arena of `total` bytes, no device."""

    def __init__(self, total, per_cell):
        self.total, self.per_cell = total, per_cell

    def slot_bytes(self, n):
        return n ** 3 * self.per_cell

    def memory(self):
        return self.total, 0, self.total

    def free_ranges(self):
        return [(0, self.total)]


@pytest.mark.timeout(30)
@pytest.mark.parametrize("admission", ["backfill", "bypass", "fcfs"])
def test_request_larger_than_the_arena_is_rejected_not_queued(admission):
    """A request whose slot exceeds the whole arena can never be admitted:
    the loop answers it with PC_ERR_NOMEM at once (it used to wait for it
    forever with nothing running)."""
    srv = sv.Server.__new__(sv.Server)
    srv.batch = {sv.TUBE: _FakeBatch(1 << 20, 1000), sv.SLAB: _FakeBatch(1 << 20, 1000)}
    srv.max_slots = 4
    srv.admission = admission
    srv.bypass_depth = 0 if admission == "fcfs" else 8
    srv.bypass_ticks = 50
    srv.plan_s = 0.0
    srv.keep_results = False
    srv.streams = {k: None for k in srv.batch}
    reqs = [sv.Request(0, sv.SLAB, 128, 10, 2.0, 1.0, 0.0, 1),
            sv.Request(1, sv.TUBE, 64, 10, 2.0, 1.0, 0.0, 2)]
    out = srv.run(reqs, realtime=False)
    assert sorted(r.rid for r in out) == [0, 1]
    assert all(r.rejected and r.status == sv.PC_ERR_NOMEM for r in out)
