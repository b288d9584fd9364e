"""Ensemble serving on one GPU: requests (3D shock tubes and premixed slabs
of different sizes and step counts) arrive over time, are admitted while
the device memory budget allows, and advance together: every tick steps
each admitted request once, with one launch per stage over all the requests
of one mechanism (continuous batching: a request that arrives while others
run joins at the next tick, a finished one leaves and frees its memory).
The host queues ticks ahead of the device and never waits for it as a
whole: completions are seen through events and the finished requests' dt
states through asynchronous snapshots.

Request fields are generated on the device from the request's own seed
(torch's counter-based CUDA generator), so a request starts from the same
bits whether it runs alone or in a mix.
"""
import bisect
import ctypes as C
import math
import time
from collections import deque

import numpy as np

from . import _lib
from .ctypes_types import Box, make_box, make_params
from .hydro import HydroContext, _stream_ptr, a4_from_tensor
from .workloads import BC, burnt_h2_air, unburnt_h2_air

PC_ERR_NOMEM = -8

check = _lib.check
TUBE, SLAB = 0, 1          # request kinds: gamma-law shock tube, premixed slab
NSPEC = {TUBE: 1, SLAB: 9}


class Batch:
    """A pc_batch_t: requests of one mechanism resident on the device."""

    def __init__(self, ctx, params, mixed, budget_bytes=0, share=None,
                 max_slots=512):
        self.ctx = ctx
        self.L = ctx.L
        self.prm = make_params(**params)
        h = C.c_void_p()
        check(self.L.pc_batch_create_v1(ctx.h, C.byref(self.prm), 1 if mixed else 0,
                                        int(budget_bytes), share.h if share else None,
                                        int(max_slots), C.byref(h)), ctx.h)
        self.h = h

    def close(self):
        if self.h:
            self.L.pc_batch_destroy_v1(self.h)
            self.h = None

    def _chk(self, rc):
        return check(rc, self.ctx.h)

    def slot_bytes(self, n):
        return self.L.pc_batch_slot_bytes_v1(self.h, make_box([0] * 3, [n - 1] * 3))

    def memory(self):
        t, u, f = C.c_longlong(), C.c_longlong(), C.c_longlong()
        self._chk(self.L.pc_batch_memory_v1(self.h, C.byref(t), C.byref(u), C.byref(f)))
        return t.value, u.value, f.value

    def free_ranges(self):
        """[(offset, bytes)] of the (possibly shared) arena, ascending."""
        n = C.c_int()
        self._chk(self.L.pc_batch_free_ranges_v1(self.h, None, None, 0, C.byref(n)))
        cap = n.value
        off, size = (C.c_longlong * max(cap, 1))(), (C.c_longlong * max(cap, 1))()
        self._chk(self.L.pc_batch_free_ranges_v1(self.h, off, size, cap, C.byref(n)))
        return [(off[i], size[i]) for i in range(min(cap, n.value))]

    def extent(self, slot):
        o, b = C.c_longlong(), C.c_longlong()
        self._chk(self.L.pc_batch_slot_extent_v1(self.h, slot, C.byref(o), C.byref(b)))
        return o.value, b.value

    def add(self, n, bclo, bchi, prim, prim_mode, stream=None):
        slot = C.c_int()
        dx = (C.c_double * 3)(*([1.0 / n] * 3))
        self._chk(self.L.pc_batch_add_v1(
            self.h, make_box([0] * 3, [n - 1] * 3), (C.c_int * 3)(*bclo),
            (C.c_int * 3)(*bchi), dx, a4_from_tensor(prim, [0, 0, 0]),
            int(prim_mode), _stream_ptr(stream), C.byref(slot)))
        return slot.value

    def set_branches(self, on=True):
        """Three streams for the three directions' stages (results unchanged)."""
        self._chk(self.L.pc_batch_set_branches_v1(self.h, 1 if on else 0))

    def step(self, slots, stream=None):
        if not slots:
            return
        arr = (C.c_int * len(slots))(*slots)
        self._chk(self.L.pc_batch_step_v1(self.h, arr, len(slots), _stream_ptr(stream)))

    def read(self, slot, n):
        out = np.empty((7 + self.ctx.nspec,) + (n + 8,) * 3)
        self._chk(self.L.pc_batch_read_v1(self.h, slot, out.ctypes.data))
        return out

    def status(self, slot):
        dt, bad, small = C.c_double(), C.c_int(), C.c_int()
        rc = self.L.pc_batch_status_v1(self.h, slot, C.byref(dt), C.byref(bad),
                                       C.byref(small))
        return rc, dt.value, bad.value, small.value

    def snapshot(self, slot, stream=None):
        """Queue a copy of the request's dt state to the host on `stream`."""
        self._chk(self.L.pc_batch_snapshot_v1(self.h, slot, _stream_ptr(stream)))

    def snapshot_status(self, slot):
        """(status, dt, bad, small) from the last snapshot, once it has
        landed (no device wait)."""
        dt, bad, small = C.c_double(), C.c_int(), C.c_int()
        rc = self.L.pc_batch_snapshot_status_v1(self.h, slot, C.byref(dt), C.byref(bad),
                                                C.byref(small))
        return rc, dt.value, bad.value, small.value

    def remove(self, slot, stream=None, wait=True):
        """Return the request's memory after waiting for `stream` (the
        current stream by default); wait=False: its work is known done."""
        ptr = _stream_ptr(stream) if wait else C.c_void_p(None)
        self._chk(self.L.pc_batch_remove_v1(self.h, slot, ptr))


class Request:
    def __init__(self, rid, kind, n, nsteps, ratio, phi, t_arrive, seed):
        self.rid, self.kind, self.n, self.nsteps = rid, kind, n, nsteps
        self.ratio, self.phi, self.t_arrive, self.seed = ratio, phi, t_arrive, seed
        self.slot = None
        self.bytes = None          # arena bytes of its slot
        self.off = None            # arena offset while it runs
        self.blocked_since = None  # tick at which it first waited at the head
        self.done = 0
        self.t_admit = self.t_done = None
        self.result = None
        self.status = None
        self.rejected = False      # larger than the whole arena: never runs
        self.ev = None             # event after its last step, while retiring

    @property
    def cell_updates(self):
        return self.n ** 3 * self.nsteps


class ArenaModel:
    """Host mirror of the batch arena's placement rule: a request takes one
    block of its bytes rounded up to 256 at the start of the free range of
    lowest offset that holds it; a freed block merges with free neighbours.
    Built from the arena's free ranges, it predicts where admissions go and
    when a request fits as running ones finish."""

    def __init__(self, ranges):
        self.fr = sorted((int(o), int(s)) for o, s in ranges if s > 0)

    @staticmethod
    def round(b):
        return (int(b) + 255) & ~255

    def copy(self):
        m = ArenaModel(())
        m.fr = list(self.fr)
        return m

    def fits(self, b):
        b = self.round(b)
        return any(s >= b for _, s in self.fr)

    def alloc(self, b):
        b = self.round(b)
        for i, (o, s) in enumerate(self.fr):
            if s >= b:
                if s > b:
                    self.fr[i] = (o + b, s - b)
                else:
                    del self.fr[i]
                return o
        return -1

    def release(self, off, b):
        b = self.round(b)
        i = bisect.bisect_left(self.fr, (off, 0))
        if i < len(self.fr) and off + b == self.fr[i][0]:
            b += self.fr[i][1]
            del self.fr[i]
        if i > 0 and self.fr[i - 1][0] + self.fr[i - 1][1] == off:
            self.fr[i - 1] = (self.fr[i - 1][0], self.fr[i - 1][1] + b)
        else:
            self.fr.insert(i, (off, b))


def ticks_until_fit(model, held, need, kind=None, slots=None):
    """Ticks until a block of `need` bytes fits, releasing the running
    requests' blocks held = [(ticks left, offset, bytes, kind)] as they
    finish; with slots ({kind: free request slots}) the request also needs
    a free slot of its kind.  None if that never happens with all of them
    released.  (Releasing only ever widens free ranges and frees slots, so
    the first fit is the answer.)"""
    free = slots[kind] if slots is not None else 1
    if free > 0 and model.fits(need):
        return 0
    m = model.copy()
    for left, off, b, k in sorted(held):
        m.release(off, b)
        free += k == kind
        if free > 0 and m.fits(need):
            return left
    return None


def backfill_plan(waiting, held, model, depth, slots=None):
    """Admissions of one tick under EASY backfilling (Lifka, 1995) over the
    first-fit arena.  waiting: queued requests in arrival order (.bytes,
    .nsteps, .kind); held: [(ticks left, offset, bytes, kind)] of the
    running requests; model: the arena now; slots: free request slots per
    kind (None: unlimited).  A request fits when its block fits and a slot
    of its kind is free.  Requests are admitted from the head while they
    fit.  A head that does not fit is given a reservation: S, the ticks until
    it fits as running requests finish.  Of the next `depth` requests, one
    that fits now is admitted when the head still fits after S ticks with it
    admitted too (it finishes within S ticks, or its block leaves the head
    room), so no admission delays the head.  model and held are updated;
    returns [(index into waiting, predicted offset)]."""
    free = dict(slots) if slots is not None else None
    fits = lambda r: (free is None or free[r.kind] > 0) and model.fits(r.bytes)

    def take(r, off):
        held.append((r.nsteps, off, ArenaModel.round(r.bytes), r.kind))
        if free is not None:
            free[r.kind] -= 1

    out = []
    i = 0
    while i < len(waiting) and fits(waiting[i]):
        r = waiting[i]
        off = model.alloc(r.bytes)
        take(r, off)
        out.append((i, off))
        i += 1
    if i == len(waiting) or depth <= 0:
        return out
    head = waiting[i]
    S = ticks_until_fit(model, held, head.bytes, head.kind, free)
    if S is None:
        return out
    for j in range(i + 1, min(len(waiting), i + 1 + depth)):
        c = waiting[j]
        if not fits(c):
            continue
        m2 = model.copy()
        off = m2.alloc(c.bytes)
        h = (c.nsteps, off, ArenaModel.round(c.bytes), c.kind)
        # one that finishes by tick S leaves the arena and the slots at S
        # as they were
        if c.nsteps > S:
            f2 = None
            if free is not None:
                f2 = dict(free)
                f2[c.kind] -= 1
            s2 = ticks_until_fit(m2, held + [h], head.bytes, head.kind, f2)
            if s2 is None or s2 > S:
                continue
        model.fr = m2.fr
        take(c, off)
        out.append((j, off))
    return out


def request_stream(count, rate, seed=20261203, n_lo=32, n_hi=128, n_med=64,
                   steps=(10, 200), ratio=(2.0, 20.0), phi=(0.5, 1.2)):
    """Poisson arrivals at `rate` per second; half tubes, half slabs; sides
    lognormal (median n_med, clipped to [n_lo, n_hi]); step counts, pressure
    ratios and equivalence ratios uniform."""
    rng = np.random.Generator(np.random.Philox(seed))
    t = np.cumsum(rng.exponential(1.0 / rate, size=count)) if rate > 0 else np.zeros(count)
    out = []
    for r in range(count):
        kind = TUBE if rng.random() < 0.5 else SLAB
        n = int(np.clip(round(n_med * math.exp(0.5 * rng.standard_normal())), n_lo, n_hi))
        out.append(Request(r, kind, n, int(rng.integers(steps[0], steps[1] + 1)),
                           float(rng.uniform(*ratio)), float(rng.uniform(*phi)),
                           float(t[r]), seed + 7919 * (r + 1)))
    return out


def request_fields(req):
    """Primitive fields of a request on the device, and its BCs and mode.
    Tube: rho, p = (1, 0.1 ratio) for x < 0.5 and (0.125, 0.1) beyond, rho
    perturbed by a uniform 1e-3 relative noise; Outflow x, periodic y/z
    (prim mode 0).  Slab: premixed H2/air at phi, T 300 -> 1800 K across a
    tanh of width 8 cells in z, p = 1 atm, velocity (0, 0, 100) cm/s plus
    N(0, 5 cm/s) noise; periodic x/y, Outflow z (prim mode 1)."""
    import torch
    n = req.n
    g = torch.Generator(device="cuda")
    g.manual_seed(req.seed)
    x = (torch.arange(n, device="cuda", dtype=torch.float64) + 0.5) / n
    if req.kind == TUBE:
        P = torch.zeros((6, n, n, n), dtype=torch.float64, device="cuda")
        left = (x <= 0.5).view(1, 1, n)
        noise = torch.rand((n, n, n), generator=g, device="cuda", dtype=torch.float64) * 2 - 1
        P[0] = torch.where(left, 1.0, 0.125) * (1.0 + 1e-3 * noise)
        P[1] = torch.where(left, 0.1 * req.ratio, 0.1).expand(n, n, n)
        P[5] = 1.0
        bc = [BC["outflow"], BC["interior"], BC["interior"]]
        return P, bc, list(bc), 0
    s = 0.5 * (1.0 + torch.tanh((x - 0.5) / (8.0 / n)))
    Yu = torch.tensor(unburnt_h2_air(req.phi), device="cuda")
    Yb = torch.tensor(burnt_h2_air(req.phi), device="cuda")
    P = torch.empty((14, n, n, n), dtype=torch.float64, device="cuda")
    P[0] = 1013250.0
    P[1] = (300.0 + 1500.0 * s).view(n, 1, 1)
    P[2:5] = torch.randn((3, n, n, n), generator=g, device="cuda", dtype=torch.float64) * 5.0
    P[4] += 100.0
    P[5:] = ((1.0 - s).view(1, n, 1, 1) * Yu.view(9, 1, 1, 1)
             + s.view(1, n, 1, 1) * Yb.view(9, 1, 1, 1))
    bc = [BC["interior"], BC["interior"], BC["outflow"]]
    return P, bc, list(bc), 1


class Server:
    """The request loop.  `budget_frac` of the device memory is one arena
    shared by the tube and slab batches; requests are admitted in arrival
    order while their state and workspace fit, and past a head that does
    not fit as `admission` says (default: backfilling that never delays the
    head).  overlap: each batch
    steps on a stream of its own (the tube and slab batches' kernels
    overlap) and runs the three directions' stages on three streams."""

    def __init__(self, mixed=True, budget_frac=0.8, budget_bytes=None,
                 max_slots=512, keep_results=False, mol=False, overlap=True,
                 bypass_depth=64, bypass_ticks=50, admission="backfill"):
        import torch
        self.ctx = {TUBE: HydroContext(1), SLAB: HydroContext(9)}
        if budget_bytes is None:
            free, total = torch.cuda.mem_get_info()
            budget_bytes = int(budget_frac * total)
            budget_bytes = min(budget_bytes, int(0.95 * free))
        self.budget = budget_bytes
        params = dict(ppm_type=1, cfl=0.3, do_mol=1 if mol else 0)
        self.batch = {}
        self.batch[SLAB] = Batch(self.ctx[SLAB], params, mixed, budget_bytes,
                                 max_slots=max_slots)
        self.batch[TUBE] = Batch(self.ctx[TUBE], params, mixed,
                                 share=self.batch[SLAB], max_slots=max_slots)
        self.keep_results = keep_results
        self.max_slots = max_slots
        # admission past a request that does not fit.  "backfill": of the
        # next bypass_depth requests, those that do not delay the head's
        # reservation (backfill_plan).  "bypass": up to bypass_depth later
        # requests that fit are admitted, until the head has waited
        # bypass_ticks ticks (then the queue holds until the head fits).
        # "fcfs": none.
        if admission not in ("fcfs", "bypass", "backfill"):
            raise ValueError(f"unknown admission policy {admission!r}")
        self.admission = admission
        self.bypass_depth = bypass_depth if admission != "fcfs" else 0
        self.bypass_ticks = bypass_ticks
        self.plan_s = 0.0          # host seconds spent planning admissions
        self.streams = {k: None for k in self.batch}
        if overlap:
            for k, b in self.batch.items():
                b.set_branches(True)
                self.streams[k] = torch.cuda.Stream()

    def close(self):
        for b in self.batch.values():
            b.close()
        for c in self.ctx.values():
            c.close()

    def _admit(self, r, t0):
        b = self.batch[r.kind]
        P, bclo, bchi, mode = request_fields(r)
        r.slot = b.add(r.n, bclo, bchi, P, mode)
        del P
        r.off, r.bytes = b.extent(r.slot)
        r.t_admit = time.perf_counter() - t0

    def _free_slots(self, holding):
        free = {k: self.max_slots for k in self.batch}
        for r in holding:
            free[r.kind] -= 1
        return free

    def _admit_backfill(self, waiting, active, retiring, t0):
        for r in waiting:
            if r.bytes is None:
                r.bytes = ArenaModel.round(self.batch[r.kind].slot_bytes(r.n))
        t = time.perf_counter()
        model = ArenaModel(self.batch[SLAB].free_ranges())
        # a retiring request's block returns once its last step has run (by
        # the next tick, as far as the plan is concerned)
        held = ([(r.nsteps - r.done, r.off, r.bytes, r.kind) for r in active] +
                [(1, r.off, r.bytes, r.kind) for r in retiring])
        plan = backfill_plan(waiting, held, model, self.bypass_depth,
                             self._free_slots(active + retiring))
        self.plan_s += time.perf_counter() - t
        for j, off in plan:
            r = waiting[j]
            self._admit(r, t0)
            if r.off != off:
                raise RuntimeError(f"arena placed request {r.rid} at {r.off}, "
                                   f"the admission plan at {off}")
            active.append(r)
        for j, _ in sorted(plan, reverse=True):
            del waiting[j]

    def run(self, requests, realtime=True, on_tick=None, ahead=2):
        """Serve `requests` (sorted by arrival).  realtime: arrivals follow
        their timestamps on the wall clock; else all are queued at t = 0.
        The host never waits for the whole device: a tick's steps are queued
        on the batches' streams and followed by an event; a request whose
        last step is queued holds its memory until that event has passed
        (its dt state comes back through an asynchronous snapshot), and at
        most `ahead` ticks are queued beyond the ones known complete.
        Returns the list of served requests with t_admit and t_done (when
        the device finished its last step) set, in seconds from the start."""
        import torch
        pending = deque(sorted(requests, key=lambda r: r.t_arrive))
        waiting = deque()
        active = []      # admitted, steps still to queue
        retiring = []    # every step queued; memory held until its event
        served = []
        inflight = deque()  # events of the queued ticks, oldest first
        t0 = time.perf_counter()
        base = None      # event at host time t_base (seconds from t0)
        t_base = 0.0
        ticks = 0

        def retire(block):
            nonlocal retiring
            left = []
            for r in retiring:
                if not (block or r.ev.query()):
                    left.append(r)
                    continue
                r.ev.synchronize()
                r.t_done = t_base + base.elapsed_time(r.ev) / 1e3
                b = self.batch[r.kind]
                if self.keep_results:
                    r.result = b.read(r.slot, r.n)
                r.status = b.snapshot_status(r.slot)[0]
                b.remove(r.slot, wait=False)
                r.ev = None
                served.append(r)
            retiring = left

        while pending or waiting or active or retiring:
            if retiring:
                retire(block=False)
            now = time.perf_counter() - t0
            while pending and (not realtime or pending[0].t_arrive <= now):
                r = pending.popleft()
                b = self.batch[r.kind]
                if ArenaModel.round(b.slot_bytes(r.n)) > b.memory()[0]:
                    # it would not fit even with every other request gone:
                    # answer it at once instead of queueing it forever
                    r.rejected = True
                    r.status = PC_ERR_NOMEM
                    r.t_admit = r.t_done = now
                    served.append(r)
                    continue
                waiting.append(r)
            if self.admission == "backfill":
                self._admit_backfill(waiting, active, retiring, t0)
            i = 0
            free = self._free_slots(active + retiring)
            while self.admission != "backfill" and i < len(waiting):
                r = waiting[i]
                b = self.batch[r.kind]
                if b.slot_bytes(r.n) > b.memory()[2] or free[r.kind] == 0:
                    if i == 0 and r.blocked_since is None:
                        r.blocked_since = ticks
                    head = waiting[0]
                    if (i + 1 > self.bypass_depth or
                            ticks - head.blocked_since >= self.bypass_ticks):
                        break
                    i += 1
                    continue
                self._admit(r, t0)
                free[r.kind] -= 1
                del waiting[i]
                active.append(r)
            if not active:
                if retiring:
                    retire(block=True)  # memory comes back only from them
                elif pending and realtime:
                    time.sleep(max(0.0, pending[0].t_arrive - (time.perf_counter() - t0)))
                continue
            while len(inflight) >= ahead:
                for ev in inflight.popleft():
                    ev.synchronize()
            # admissions ran on the current stream (pc_batch_add_v1 returns
            # once they have); each tick's steps after them
            cur = torch.cuda.current_stream()
            if base is None:
                base = torch.cuda.Event(enable_timing=True)
                base.record(cur)
                t_base = time.perf_counter() - t0
            evs = {}
            for kind, b in self.batch.items():
                mine = [r for r in active if r.kind == kind]
                if not mine:
                    continue
                st = self.streams[kind] or cur
                if st is not cur:
                    st.wait_stream(cur)
                b.step([r.slot for r in mine], st)
                for r in mine:
                    r.done += 1
                    if r.done == r.nsteps:
                        b.snapshot(r.slot, st)
                ev = torch.cuda.Event(enable_timing=True)
                ev.record(st)
                evs[kind] = ev
            inflight.append(list(evs.values()))
            ticks += 1
            for r in active:
                if r.done == r.nsteps:
                    r.ev = evs[r.kind]
                    retiring.append(r)
            active = [r for r in active if r.done < r.nsteps]
            if on_tick:
                on_tick(ticks, active + retiring)
        self.wall = time.perf_counter() - t0
        self.ticks = ticks
        return served


def summarize(served, wall):
    rejected = sum(1 for r in served if r.rejected)
    served = [r for r in served if not r.rejected]
    lat = np.array([r.t_done - r.t_arrive for r in served])
    cells = sum(r.cell_updates for r in served)
    return {"requests": len(served), "rejected": rejected, "wall_s": wall,
            "mcell_updates_per_s": cells / wall / 1e6,
            "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                           "p99": float(np.percentile(lat, 99) * 1e3),
                           "mean": float(lat.mean() * 1e3)},
            "queue_ms": {"p50": float(np.percentile([r.t_admit - r.t_arrive for r in served], 50) * 1e3),
                         "p99": float(np.percentile([r.t_admit - r.t_arrive for r in served], 99) * 1e3)}}
