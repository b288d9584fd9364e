#!/usr/bin/env python3
"""Godunov hydro step throughput on one MI355X.

Default workload: premixed H2/air, 9 species (LiDryer ideal-gas mixture),
hydro only, 128^3 cells as 64 boxes of 32^3 with 4 ghost cells, PPM, the
reference's approximate Riemann solver, CFL 0.3 (the FP64 premixed H2/air
workload the throughput metric is quoted on).  One "step" = ghost fill + the full hydro step
(ctoprim -> CTU fluxes -> flux divergence -> S_new -> internal-energy reset
and e->T) + the next dt, for every box of the level, on the GPU.

Prints ONE JSON line.  `value` = valid cells x timed steps / seconds / 1e6
(Mcell-updates/s) with the state resident in HBM.  `roofline` is the
dominant kernel against HBM bandwidth, from its stage-compulsory bytes and
its HIP-event duration measured here, with `traffic` (HBM bytes per launch
from calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE, profiles/); `hbm_frac_step`
is the whole step against HBM on the reference's algorithmic bytes
(2 * NVAR * 8 per cell-update) and `fp64_step` the whole step against FP64
peak on the counted FP64 operations of one step (same profile).  `cpu_baseline` is the CPU oracle (this repository's C++
restatement of the reference path, OpenMP over boxes) on one step of the
same level, timed on this host.

`--mol` runs the method-of-lines step instead (do_mol = 1: two Heun stages
of ghost fill, primitives, PLM-MOL fluxes and update per step; `--laxf`
for the Lax-Friedrichs flux); the oracle baseline then runs the same step.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # B/s, MI355X HBM3E
FP64_PEAK = 78.6e12        # FLOP/s, MI355X FP64 vector
METRIC = ("Mcell-updates/s per hydro step + % HBM roofline; "
          "p99 ensemble-solve latency")
# committed rocprofv3 summaries of the 128^3 premixed level step, by
# (method of lines, mixed precision)
PROFILE_SUMMARIES = {
    (False, False): "premixed128_step_kernels.md",
    (False, True): "premixed128_mixed_step_kernels.md",
    (True, False): "premixed128_mol_step_kernels.md",
    (True, True): "premixed128_mol_mixed_step_kernels.md",
}
FP32_PEAK = 157.3e12
# launches per step of the kernels the MOL step runs in both Heun stages
MOL_LAUNCHES = {"ctoprim": 2, "mol_x": 2, "mol_y": 2, "mol_z": 2}


def stage_bytes(nspec, boxes, mol=False, fused=True):
    """Compulsory bytes of one launch of each kernel over the level: every
    input element read once and every output element written once on the
    kernel's region (stencil re-reads are not counted).  fused: the PPM
    trace launch also forms the first fluxes and each cell's transverse
    correction (k_trace_flux1_lds); otherwise k_flux1 writes the face fluxes
    and k_tcorr turns them into the corrections."""
    Q, N = 8 + nspec, 7 + nspec
    # the Godunov level step stores a traced state and a transverse
    # correction without rho e and the energy flux difference (TS values:
    # rho, u, v, w, p, Y; the correction's pressure term in place of the
    # energy), a transverse flux without the energy flux (FS: rho, momenta,
    # p_gd, u_gd, Y), and final fluxes without the UTEMP and UEINT fluxes and
    # without qec (the update resets UEINT)
    TS, FS = 5 + nspec, 6 + nspec
    tot = {}

    def add(k, units, per):
        tot[k] = tot.get(k, 0) + units * per * 8

    for lo, hi in boxes:
        n = [hi[d] - lo[d] + 1 for d in range(3)]
        g4 = (n[0] + 8) * (n[1] + 8) * (n[2] + 8)
        v = n[0] * n[1] * n[2]
        if mol:
            # primitives on grow(bx,2) (ghost values read at their source
            # cells); per d-face the primitives the slopes read (rho, u, v,
            # w, p, Y, c) of one cell and the NVAR fluxes written; the first
            # update reads U^n and the three fluxes and writes U* to the stage
            # array, the second reads U*, U^n and the fluxes and writes U
            add("ctoprim", (n[0] + 4) * (n[1] + 4) * (n[2] + 4), N + Q + 6)
            for d, a in enumerate("xyz"):
                f = list(n)
                f[d] += 1
                add("mol_" + a, f[0] * f[1] * f[2], 6 + nspec + N)
            add("mol_update1", v, 5 * N)
            add("mol_update2", v, 6 * N)
            add("dt", 0, 0)
            continue
        g1 = (n[0] + 2) * (n[1] + 2) * (n[2] + 2)
        # the primitives read every cell's state once (a ghost cell's at its
        # source) and store the first ghost layer of U back (no ghost fill)
        add("ctoprim", g4, N + Q + 6)
        add("ctoprim", g1 - v, N)
        # divu on the nodes (4 velocity reads, one write) and the flattening
        # coefficient on grow(bx,1) (p and the velocity read, one write)
        add("divu", (n[0] + 2) * (n[1] + 2) * (n[2] + 2), 6)
        for d, a in enumerate("xyz"):
            t = [x + 2 for x in n]
            add("trace_" + a, g1, TS + 2 * TS)
            f1 = list(t)
            f1[d] = n[d] + 1
            # cells [L, H] along d (transverse [L-1, H+1]) get a correction
            # of TS values (rho, m, rho Y and the pressure term) from Gamma1
            # and their two faces' fluxes
            c1 = list(t)
            c1[d] = n[d]
            if fused:  # the faces' states come from the trace, not from HBM
                add("trace_" + a, f1[0] * f1[1] * f1[2], 1)
                add("trace_" + a, c1[0] * c1[1] * c1[2], 1 + TS)
            else:
                add("flux1_" + a, f1[0] * f1[1] * f1[2], 2 * TS + 1 + FS)
                add("flux1_" + a, c1[0] * c1[1] * c1[2], FS + 1 + TS)
            fin = list(n)
            fin[d] = n[d] + 1
            # two traces, the two transverse fluxes (each once), U at one
            # cell (artificial viscosity; neither UTEMP nor UEINT), the flux
            # written, c and Gamma1 at two cells and div u
            add("final_" + a, fin[0] * fin[1] * fin[2], 2 * TS + 2 * FS + 2 * (N - 2) + 5)
        for D, name in [(0, "xy"), (1, "yx"), (2, "zx")]:
            # one launch per direction for both (d|t): on the d-faces with
            # both transverse indices in [L-1, H+1] the face's traces and c
            # are read once; each (d|t) reads the t-correction of one cell
            # per face and writes its flux where t lies in [L, H]
            T0, T1 = [t for t in range(3) if t != D]
            r = [x + 2 for x in n]
            r[D] = n[D] + 1
            add("trans_" + name, r[0] * r[1] * r[2], 2 * TS + 1)
            for T in (T0, T1):
                e = list(r)
                e[T] = n[T]
                add("trans_" + name, e[0] * e[1] * e[2], TS + FS)
        # U read (not UEINT) and written, three face fluxes per cell
        add("update", v, (N - 1) + N + 3 * (N - 2))
        add("dt", 0, 0)
    return tot


def profiled_traffic(summary, stage):
    """HBM bytes per launch of `stage` from a committed rocprofv3 summary
    (calibrated FETCH_SIZE + WRITE_SIZE, column `HBM MB`), or None."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    col = None
    for line in open(path):
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if "HBM MB" in cols and "stage" in cols:
            col = cols.index("HBM MB")
        elif col is not None and len(cols) > col and cols[1] == stage:
            return float(cols[col]) * 1e6
    return None


def profiled_columns(summary, stage, names):
    """Columns `names` of the row of `stage` in a committed rocprofv3
    summary, as floats (None where absent)."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return {k: None for k in names}
    head = None
    for line in open(path):
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if "HBM MB" in cols and "stage" in cols:
            head = cols
        elif head is not None and len(cols) == len(head) and cols[1] == stage:
            out = {}
            for k in names:
                try:
                    out[k] = float(cols[head.index(k)])
                except (ValueError, IndexError):
                    out[k] = None
            return out
    return {k: None for k in names}


def profiled_step_traffic(summary):
    """Counted HBM bytes of one level step from the same summary, or None."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("Sum over one step") and "HBM traffic" in line:
            tok = line.split("HBM traffic ")[1].split(" GB")[0]
            return float(tok) * 1e9
    return None


def profiled_step_flops(summary, kind="FP64"):
    """FP64 (or FP32) operations of one level step from the same summary
    (counted SQ_INSTS_VALU_*_F64 / _F32 x 64 lanes), or None."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("Sum over one step"):
            tok = line.split(kind + " ")[1].split(" GFLOP")[0]
            return float(tok) * 1e9
    return None


def oracle_norms(w, gpu_fabs, host_fabs):
    """L1/Linf relative errors in rho and T of device box states against the
    oracle's (valid cells)."""
    import numpy as np
    acc = {q: [0.0, 0.0, 0.0, 0.0] for q in ("rho", "T")}
    for g, h in zip(gpu_fabs, host_fabs):
        for q, c in (("rho", 0), ("T", 6)):
            a = g[c, 4:-4, 4:-4, 4:-4]
            b = h[c, 4:-4, 4:-4, 4:-4]
            d = np.abs(a - b)
            v = acc[q]
            v[0] += d.sum()
            v[1] += np.abs(b).sum()
            v[2] = max(v[2], d.max())
            v[3] = max(v[3], np.abs(b).max())
    return {q: {"L1": v[0] / v[1], "Linf": v[2] / v[3]} for q, v in acc.items()}


def precision_check(ctx, w, steps, with_oracle):
    """Fresh FP64 and mixed levels from the initial state, stepped in
    lockstep to each checkpoint: error norms of the mixed state against the
    FP64 state (and, at the first checkpoint when it is step 1, against the
    CPU oracle), plus each level's wall time per step."""
    import torch
    from particleclem_amd.hydro import DeviceLevel
    from particleclem_amd.precision import error_norms
    l64 = DeviceLevel.from_workload(ctx, w)
    l32 = DeviceLevel.from_workload(ctx, w)
    l32.set_precision(True)
    for x in (l64, l32):
        x.init_dt()
    torch.cuda.synchronize()
    t64 = t32 = 0.0
    done = 0
    res = {}
    for s in steps:
        n = s - done
        t0 = time.perf_counter()
        l64.step(n)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        l32.step(n)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t64 += t1 - t0
        t32 += t2 - t1
        done = s
        e = error_norms(l32, l64, w.nspec, len(w.boxes))
        ys = [e[f"Y{k}"] for k in range(w.nspec)] if w.nspec > 1 else []
        r = {"rho": e["rho"], "T": e["T"]}
        if ys:
            r["Y_max"] = {"L1": max(y["L1"] for y in ys),
                          "Linf": max(y["Linf"] for y in ys)}
            r["Y"] = {f"Y{k}": e[f"Y{k}"] for k in range(w.nspec)}
        res[str(s)] = r
        if with_oracle and s == 1:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import oracle_bind as ob
            H = ob.HostLevel(w.nspec, w.boxes, w.domlo, w.domhi, w.bclo,
                             w.bchi, w.dx, periodic_self=w.periodic_self)
            H.set_prims(w)
            prm = ob.params(**w.params)
            H.step(prm, 1, H.estdt(prm) * prm.init_shrink, 0)
            g = [l32.read_box(b) for b in range(len(w.boxes))]
            r["vs_cpu_oracle"] = oracle_norms(w, g, [H.fab(b) for b in range(len(w.boxes))])
    out = {"reference": "FP64 step on the GPU (bitwise equal to the CPU "
                        "oracle in the parity tests); dt from each run's own "
                        "CFL estimate",
           "errors": res, "fp64_ms_per_step": t64 / done * 1e3,
           "mixed_ms_per_step": t32 / done * 1e3, "mixed_speedup": t64 / t32}
    for x in (l64, l32):
        x.close()
    return out


def serve(args):
    """Ensemble serving: a seeded stream of shock-tube / premixed-slab
    requests with Poisson arrivals, served by continuous batching under an
    80% HBM budget; value = aggregate Mcell-updates/s over the run, with the
    p50/p99 latency (arrival to result ready, queueing included)."""
    from particleclem_amd import serving as sv
    reqs = sv.request_stream(args.requests, args.rate)
    srv = sv.Server(mixed=(args.precision == "mixed"), mol=args.mol,
                    bypass_depth=args.bypass_depth, admission=args.admission)
    warm = sv.request_stream(4, 0, seed=1, n_lo=16, n_hi=16, n_med=16, steps=(3, 3))
    srv.run(warm, realtime=False)        # kernel loading, allocator warm-up
    served = srv.run(reqs, realtime=True)
    summ = sv.summarize(served, srv.wall)
    bad = sum(1 for r in served if r.status != 0)
    budget = srv.budget
    srv_admission, plan_s = srv.admission, srv.plan_s
    srv.close()
    cells = sum(r.cell_updates for r in served)
    out = {
        "metric": METRIC, "value": summ["mcell_updates_per_s"],
        "unit": "Mcell-updates/s", "n_gpus": 1, "steps": None, "warmup": 4,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64 (mixed)",
        "data": "synthetic",
        "config": {"workload": f"serving: {args.requests} requests, Poisson "
                               f"arrivals at {args.rate}/s, half 3D shock tubes "
                               f"(gamma-law, pressure ratio 2-20) and half "
                               f"premixed H2/air slabs (9 species, phi 0.5-1.2), "
                               f"sides lognormal in [32, 128] (median 64), 10-200 "
                               f"steps, {'MOL' if args.mol else 'PPM'}, CFL 0.3, "
                               f"continuous batching under an 80% HBM budget "
                               f"({budget / 2**30:.0f} GiB), admission "
                               f"{srv_admission} (depth {args.bypass_depth})",
                   "requests": len(served), "cell_updates": cells,
                   "failed": bad},
        "latency_ms": summ["latency_ms"], "queue_ms": summ["queue_ms"],
        "wall_s": summ["wall_s"], "ticks": srv.ticks,
        "admission_plan_s": plan_s,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def dropin(args):
    """The drop-in form: PeleC's construct_hydro_source calling the library
    once per box (pc_hydro_source_box_v1: primitives, trace, Riemann solves,
    transverse corrections, flux adjustment and dU/dt of one box, with its
    ghost cells already filled, caller-owned FABs for the state, hydro
    source, fluxes, areas and volume), over the 64 boxes of 32^3 of the
    premixed 128^3 level per evaluation, the boxes spread over --streams
    streams in rotation as AMReX's MFIter does on GPUs (16 by default).
    value = valid cells x evaluations / wall time; no ghost fill, update or
    time step (the caller's part).  host_issue_ms_per_step: host time spent
    in the calls (they return once their launches are queued)."""
    import ctypes as C
    import torch
    from particleclem_amd import _lib, workloads
    from particleclem_amd.ctypes_types import Array4, make_box, make_params, null_a4
    from particleclem_amd.hydro import (DeviceLevel, HydroContext, _d3, _i3,
                                        _stream_ptr, a4_from_tensor)
    n = args.n or 128
    w = workloads.premixed(n=n, box=args.box, nsteps=1)
    ctx = HydroContext(w.nspec)
    lev = DeviceLevel.from_workload(ctx, w)
    lev.init_dt()
    lev.fill_ghosts()
    dt = lev.get_dt()
    nv = w.nvar
    dx = w.dx
    prm = make_params(**w.params)
    streams = [torch.cuda.Stream() for _ in range(max(1, args.streams))]
    L = _lib.lib()
    calls, keep = [], []
    for b, (lo, hi) in enumerate(w.boxes):
        U = torch.empty(lev.fab_shape(b), dtype=torch.float64, device="cuda")
        _lib.check(L.pc_level_read_box_v1(lev.h, b, C.c_void_p(U.data_ptr())))
        shp = tuple(hi[d] - lo[d] + 1 for d in (2, 1, 0))
        hs = torch.zeros((nv,) + shp, dtype=torch.float64, device="cuda")
        fl, ar = [], []
        for d in range(3):
            fs = list(shp)
            fs[2 - d] += 1
            fl.append(torch.zeros((nv,) + tuple(fs), dtype=torch.float64, device="cuda"))
            ar.append(torch.full((1,) + tuple(fs), [dx[1] * dx[2], dx[2] * dx[0],
                                                    dx[0] * dx[1]][d],
                                 dtype=torch.float64, device="cuda"))
        vol = torch.full((1,) + shp, dx[0] * dx[1] * dx[2], dtype=torch.float64,
                         device="cuda")
        g4 = [x - 4 for x in lo]
        keep += [U, hs, vol] + fl + ar
        calls.append((ctx.h, make_box(lo, hi), _i3(w.domlo), _i3(w.domhi),
                      _i3(w.bclo), _i3(w.bchi), a4_from_tensor(U, g4), null_a4(),
                      a4_from_tensor(hs, lo), _d3(dx), float(dt), C.byref(prm),
                      (Array4 * 3)(*[a4_from_tensor(f, lo) for f in fl]),
                      (Array4 * 3)(*[a4_from_tensor(a, lo) for a in ar]),
                      a4_from_tensor(vol, lo), _stream_ptr(streams[b % len(streams)])))
    lev.close()
    fn = L.pc_hydro_source_box_v1

    def evaluation():
        for c in calls:
            rc = fn(*c)
            if rc != 0:
                _lib.check(rc, ctx.h)

    for _ in range(args.warmup):
        evaluation()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    issue = 0.0
    for _ in range(args.steps):
        ti = time.perf_counter()
        evaluation()
        issue += time.perf_counter() - ti
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    cells = w.valid_cells()
    out = {
        "metric": METRIC, "value": cells * args.steps / sec / 1e6,
        "unit": "Mcell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": sec / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"drop-in: pc_hydro_source_box_v1 once per box, "
                               f"{len(w.boxes)} boxes of {args.box}^3 (premixed "
                               f"H2/air {n}^3, 9 species, PPM) per evaluation; "
                               f"hydro source and fluxes only (no ghost fill, "
                               f"update or dt), boxes over {len(streams)} streams",
                   "valid_cells": cells, "boxes": len(w.boxes), "calls_per_step": len(calls)},
        "host_issue_ms_per_step": issue / args.steps * 1e3,
        "roofline": None, "cpu_baseline": None,
    }
    ctx.close()
    print(json.dumps(out))


def forms(args, ctx, w, fp64_ms):
    """The forms beside the plain FP64 level step, measured in the same run:
    mixed precision on the same level (timed as the FP64 step is: branch
    streams, warm-up steps, then the same number of timed steps; error norms
    from FP64 and mixed levels stepped in lockstep), the method-of-lines step
    on the same level (FP64 and mixed, timed alike), the ragged 300-box level
    (FP64, branch streams) and a serving run of the ensemble generator (sides
    32-128, 10-200 steps, mixed precision)."""
    import torch
    from particleclem_amd import workloads
    from particleclem_amd.hydro import DeviceLevel, HydroContext
    out = {}
    t0 = time.perf_counter()
    lm = DeviceLevel.from_workload(ctx, w)
    lm.set_precision(True)
    if args.branches:
        lm.set_branches(True)
    lm.init_dt()
    lm.step(args.warmup)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    lm.step(args.steps)
    torch.cuda.synchronize()
    mixed_ms = (time.perf_counter() - ts) / args.steps * 1e3
    mrc = lm.status()[0]
    lm.close()
    pc = precision_check(ctx, w, [1, 10, 50], not args.no_cpu_baseline)
    cells = w.valid_cells()
    out["mixed"] = {
        "workload": "the same premixed 128^3 level, FP32 reconstruction, "
                    "limiters, Riemann solves and intermediates, FP64 EOS "
                    "inversion, internal energy and update; timed as the FP64 "
                    f"line ({args.warmup} warm-up and {args.steps} timed steps"
                    f"{', branch streams' if args.branches else ''}); errors from "
                    "FP64 and mixed levels stepped in lockstep from the initial "
                    "state, each with its own CFL dt, compared at steps 1, 10, 50",
        "fp64_ms_per_step": fp64_ms, "mixed_ms_per_step": mixed_ms,
        "mixed_value": cells / (mixed_ms * 1e-3) / 1e6,
        "speedup": fp64_ms / mixed_ms, "status": mrc,
        "lockstep_ms_per_step": {"fp64": pc["fp64_ms_per_step"],
                                 "mixed": pc["mixed_ms_per_step"],
                                 "note": "one stream, host-synchronised at "
                                         "each checkpoint, first steps included"},
        "errors_vs_fp64": pc["errors"], "wall_s": time.perf_counter() - t0}
    # the method-of-lines step on the same level, FP64 and mixed
    t0 = time.perf_counter()
    mol = {}
    for name, mixed in (("fp64", False), ("mixed", True)):
        lv = DeviceLevel.from_workload(ctx, w, do_mol=1)
        lv.set_precision(mixed)
        if args.branches:
            lv.set_branches(True)
        lv.init_dt()
        lv.step(args.warmup)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        lv.step(args.steps)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - ts) / args.steps * 1e3
        mol[name] = {"ms_per_step": ms, "value": cells / (ms * 1e-3) / 1e6,
                     "status": lv.status()[0]}
        lv.close()
    out["mol"] = {
        "workload": "the same premixed 128^3 level stepped by the method-of-lines "
                    "step (do_mol = 1: two Heun stages of piecewise-linear face "
                    "states and Riemann fluxes), FP64 and mixed precision, timed "
                    "as the FP64 line",
        "fp64": mol["fp64"], "mixed": mol["mixed"],
        "speedup": mol["fp64"]["ms_per_step"] / mol["mixed"]["ms_per_step"],
        "wall_s": time.perf_counter() - t0}
    # ragged level (c2)
    t0 = time.perf_counter()
    wr = workloads.ragged(nboxes=300, nsteps=args.forms_steps)
    cr = HydroContext(wr.nspec)
    lr = DeviceLevel.from_workload(cr, wr)
    lr.set_branches(True)
    lr.init_dt()
    lr.step(2)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    lr.step(args.forms_steps)
    torch.cuda.synchronize()
    sec = time.perf_counter() - ts
    rc, bad, small = lr.status()
    rcells = wr.valid_cells()
    lr.close()
    cr.close()
    out["ragged"] = {
        "workload": "one AMR level as a ragged table of 300 boxes, sides from "
                    "{8,16,24,32,48,64} with probabilities (0.1,0.2,0.2,0.25,"
                    "0.15,0.1), each its own periodic domain, premixed 9-species "
                    "generator, FP64, PPM, one launch per stage over all boxes, "
                    "branch streams",
        "valid_cells": rcells, "steps": args.forms_steps,
        "ms_per_step": sec / args.forms_steps * 1e3,
        "value": rcells * args.forms_steps / sec / 1e6,
        "hbm_frac_step": rcells * 2 * wr.nvar * 8 / (sec / args.forms_steps) / HBM_PEAK,
        "status": rc, "wall_s": time.perf_counter() - t0}
    # serving (c4)
    from particleclem_amd import serving as sv
    t0 = time.perf_counter()
    reqs = sv.request_stream(args.serving_requests, args.serving_rate)
    srv = sv.Server(mixed=True)
    warm = sv.request_stream(4, 0, seed=1, n_lo=16, n_hi=16, n_med=16, steps=(3, 3))
    srv.run(warm, realtime=False)
    served = srv.run(reqs, realtime=True)
    summ = sv.summarize(served, srv.wall)
    budget = srv.budget
    srv.close()
    out["serving"] = {
        "workload": f"{args.serving_requests} requests of the ensemble generator "
                    f"(half 3D shock tubes, gamma-law, pressure ratio 2-20; half "
                    f"premixed H2/air slabs, 9 species, phi 0.5-1.2; sides "
                    f"lognormal in [32, 128], median 64; 10-200 steps), Poisson "
                    f"arrivals at {args.serving_rate}/s, mixed precision, PPM, "
                    f"continuous batching under an 80% HBM budget "
                    f"({budget / 2**30:.0f} GiB), backfilling admission",
        "requests": summ["requests"], "rejected": summ["rejected"],
        "failed": sum(1 for r in served if r.status not in (0, None) and not r.rejected),
        "value": summ["mcell_updates_per_s"], "unit": "Mcell-updates/s",
        "latency_ms": summ["latency_ms"], "queue_ms": summ["queue_ms"],
        "wall_s": time.perf_counter() - t0}
    return out


def host_cpu():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


_NATIVE_CHECK = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[2])
sys.path.insert(0, sys.argv[2] + "/tests")
from particleclem_amd import workloads as wl
from parity_util import run_oracle
w = wl.premixed(n=16, box=8, nsteps=2)
r, dts = run_oracle(w, 2)
np.save(sys.argv[1], np.concatenate([x.ravel() for x in r] + [np.array(dts)]))
"""


def native_oracle():
    """Build the oracle for this host (-O3 -march=native, oracle/Makefile
    `native`) and check it bitwise against the portable build on a small
    premixed level (two steps); returns (path or None, note)."""
    import subprocess
    import tempfile
    import numpy as np
    odir = os.path.join(ROOT, "oracle")
    so = os.path.join(odir, "liboracle_native.so")
    r = subprocess.run(["make", "-s", "-C", odir, "native"], capture_output=True, text=True)
    if r.returncode != 0 or not os.path.exists(so):
        return None, "native build failed: " + r.stderr[-200:]
    with tempfile.TemporaryDirectory() as td:
        script = os.path.join(td, "check.py")
        open(script, "w").write(_NATIVE_CHECK)
        outs = []
        for lib in (os.path.join(odir, "liboracle.so"), so):
            out = os.path.join(td, "o%d" % len(outs))
            env = dict(os.environ, PC_ORACLE_SO=lib)
            if subprocess.run([sys.executable, script, out, ROOT], env=env,
                              capture_output=True).returncode != 0:
                return None, "native check failed to run"
            outs.append(np.load(out + ".npy"))
    if not np.array_equal(outs[0], outs[1]):
        return None, "native build differs from the portable one: not used"
    return so, "-O3 -march=native, bitwise equal to the portable build on a 2-step premixed check"


def cpu_baseline(w, gpu_fabs, dt, steps, threads, reps=3):
    """The oracle stepping the same level from the same state and dt:
    median over `reps` timings of `steps` steps each."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    H = ob.HostLevel(w.nspec, w.boxes, w.domlo, w.domhi, w.bclo, w.bchi, w.dx,
                     periodic_self=w.periodic_self)
    prm = ob.params(**w.params)
    times = []
    for _ in range(reps):
        for b, f in enumerate(gpu_fabs):
            H.fab(b)[...] = f
        t0 = time.perf_counter()
        H.step(prm, steps, dt, threads)
        times.append(time.perf_counter() - t0)
    t = sorted(times)[len(times) // 2]
    return w.valid_cells() * steps / t / 1e6, t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="premixed",
                    choices=["premixed", "sod", "ragged", "serving", "dropin"])
    ap.add_argument("--rate", type=float, default=100.0,
                    help="serving: Poisson arrival rate (requests/s)")
    ap.add_argument("--requests", type=int, default=300,
                    help="serving: requests in the stream")
    ap.add_argument("--nboxes", type=int, default=300,
                    help="ragged workload: boxes in the table")
    ap.add_argument("--ppm-type", type=int, default=1)
    ap.add_argument("--mol", action="store_true",
                    help="method-of-lines step (do_mol = 1) instead of CTU")
    ap.add_argument("--laxf", action="store_true",
                    help="MOL with the Lax-Friedrichs flux (use_laxf_flux = 1)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick every stage's launch shape on the level first")
    ap.add_argument("--bypass-depth", type=int, default=64,
                    help="serving: requests examined for admission past a "
                         "head request that does not fit")
    ap.add_argument("--admission", default="backfill",
                    choices=["fcfs", "bypass", "backfill"],
                    help="serving: admission past a head that does not fit; "
                         "backfill admits only requests that do not delay "
                         "the head, bypass any that fit (for a bounded wait "
                         "of the head), fcfs none")
    ap.add_argument("--streams", type=int, default=16,
                    help="dropin: streams the per-box calls rotate over "
                         "(AMReX: amrex.max_gpu_streams, default 4)")
    ap.add_argument("--branches", type=int, default=1, choices=[0, 1],
                    help="1: the three directions' stages on three streams in "
                         "the timed steps (kernel timing steps run on one)")
    ap.add_argument("--graph", type=int, default=0,
                    help="replay steps from a hipGraph of this many steps")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="mixed: FP32 reconstruction/limiters/Riemann and "
                         "intermediates, FP64 EOS inversion and update")
    ap.add_argument("--error-steps", default="",
                    help="comma list of steps: afterwards step fresh FP64 and "
                         "mixed levels from the initial state in lockstep and "
                         "report error norms (and both step times) there")
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--box", type=int, default=32)
    ap.add_argument("--kernel-steps", type=int, default=3)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="premixed default workload: skip the forms measured "
                         "after the main line (mixed precision, the "
                         "method-of-lines step, the ragged 300-box level, a "
                         "serving run)")
    ap.add_argument("--forms-steps", type=int, default=10,
                    help="timed steps of the ragged-level form")
    ap.add_argument("--serving-requests", type=int, default=200,
                    help="requests of the serving form (the ensemble "
                         "generator: sides 32-128, 10-200 steps)")
    ap.add_argument("--serving-rate", type=float, default=50.0,
                    help="Poisson arrival rate of the serving form (req/s)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one plain device copy of this size first (the "
                         "reference kernel for scaling HBM counters)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build runs the hydro path on a "
              "single MI355X (multi-GPU sharding is design-only); use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from particleclem_amd import workloads
    from particleclem_amd.hydro import DeviceLevel, HydroContext
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.workload == "serving":
        return serve(args)
    if args.workload == "dropin":
        return dropin(args)

    if args.workload == "premixed":
        n = args.n or 128
        w = workloads.premixed(n=n, box=args.box, nsteps=args.steps)
        desc = (f"premixed H2/air, 9 species (LiDryer ideal-gas mixture), hydro "
                f"only, {n}^3 as {len(w.boxes)} boxes of {args.box}^3 + 4 ghosts, "
                f"PPM, approximate Riemann solver, CFL 0.3, x/y periodic, "
                f"z outflow")
    elif args.workload == "ragged":
        n = None
        w = workloads.ragged(nboxes=args.nboxes, nsteps=args.steps,
                             ppm_type=args.ppm_type)
        desc = (f"ragged table of {len(w.boxes)} boxes, sides from "
                f"{{8,16,24,32,48,64}} with probabilities (0.1,0.2,0.2,0.25,"
                f"0.15,0.1), each its own periodic domain, premixed H2/air "
                f"9-species generator, {'PPM' if args.ppm_type else 'PLM'}, "
                f"approximate Riemann solver, CFL 0.3")
    else:
        n = args.n or 64
        w = workloads.sod(n=n, nsteps=args.steps)
        desc = (f"3D Sod, gamma-law gas, {n}^3 single box + 4 ghosts, PPM, "
                f"approximate Riemann solver, CFL 0.3")

    if args.mol:
        w.params.update(do_mol=1, use_laxf_flux=int(args.laxf))
        desc = desc.replace("PPM, " if "PPM, " in desc else "PLM, ", "")
        desc += (", method of lines (do_mol: PLM-MOL face states, two-stage "
                 f"Heun), {'Lax-Friedrichs flux' if args.laxf else 'approximate Riemann solver'}")
        desc = desc.replace(", approximate Riemann solver, CFL", ", CFL")

    if args.copy_calibration_mib:
        # two copies of known size (read N bytes, write N bytes) whose HBM
        # counter readings calibrate FETCH_SIZE / WRITE_SIZE: the runtime's
        # blit kernel and an elementwise multiply by one
        nel = args.copy_calibration_mib * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    ctx = HydroContext(w.nspec)
    lev = DeviceLevel.from_workload(ctx, w)
    lev.set_precision(args.precision == "mixed")
    if args.branches:
        lev.set_branches(True)
    lev.init_dt()
    torch.cuda.synchronize()
    cells = w.valid_cells()

    # warm-up (JIT-free: kernels are precompiled; this settles clocks/caches)
    lev.step(args.warmup)
    torch.cuda.synchronize()
    plans = None
    if args.autotune:
        lev.autotune(reps=3)
        plans = {k: [v[0] if isinstance(v[0], str) else list(v[0]), v[1], v[2],
                     round(v[3], 4)] for k, v in lev.plans().items()}
    if args.graph:
        lev.set_graph(args.graph)
        lev.step(args.graph)  # capture + one replay outside the timed region
        torch.cuda.synchronize()
    fabs0 = None
    dt_cpu = None
    if not args.no_cpu_baseline:
        fabs0 = [lev.read_box(b) for b in range(len(w.boxes))]
        dt_cpu = lev.get_dt()

    # timed region: K full steps, state resident in HBM
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lev.step(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    rc, bad, small = lev.status()
    if rc != 0:
        print(f"state check failed: status {rc} (nan={bad}, small_dt={small})",
              file=sys.stderr)
        sys.exit(1)
    sec = t1 - t0
    ms_step = sec / args.steps * 1e3
    value = cells * args.steps / sec / 1e6

    # per-kernel device times (events around every launch, separate steps)
    lev.enable_timing(True)
    lev.step(args.kernel_steps)
    torch.cuda.synchronize()
    kt = lev.kernel_times()
    lev.enable_timing(False)
    fused = (not args.mol) and w.params.get("ppm_type", 1) == 1 and kt.get("flux1_x", 1.0) < 0.02
    kbytes = stage_bytes(w.nspec, w.boxes, mol=args.mol, fused=fused)
    launches = MOL_LAUNCHES if args.mol else {}
    dom = max(kt, key=lambda k: kt[k])
    dom_ms = kt[dom] / launches.get(dom, 1)   # average launch duration
    achieved = kbytes[dom] / (dom_ms * 1e-3)
    summary = PROFILE_SUMMARIES[(bool(args.mol), args.precision == "mixed")]
    traffic = None
    if (args.workload == "premixed" and n == 128 and args.box == 32
            and not args.autotune and not args.laxf):
        traffic = profiled_traffic(summary, dom)
    kname = ("k_mol_update" if dom.startswith("mol_update") else
             ("k_mol_flux_lds" if args.precision == "fp64" else "k_mol_flux")
             if dom.startswith("mol_") else
             "k_trace_flux1_lds" if dom.startswith("trace_") and kt.get("flux1_x", 1) < 0.02
             else "k_" + dom.split("_")[0])
    # the dominant kernel against HBM on its own stage bytes (every input
    # element of its region read once, every output written once)
    roof_kernel = {"bound": "hbm", "kernel": f"{kname} ({dom})",
                   "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                   "frac": achieved / HBM_PEAK, "traffic": traffic,
                   "traffic_source": (f"profiles/{summary}" if traffic else None),
                   "bytes_per_launch": kbytes[dom], "ms_per_launch": dom_ms,
                   "launches_per_step": launches.get(dom, 1)}
    if traffic is not None:
        # the clock the kernel ran at and its FP64 vector pipe occupancy
        # (FP64 instructions x 4 cycles over the 1024 SIMDs' cycles)
        pc = profiled_columns(summary, dom, ["clock GHz", "FP64 pipe"])
        if pc["clock GHz"] is not None:
            roof_kernel["clock_ghz"] = pc["clock GHz"]
            roof_kernel["fp64_pipe"] = pc["FP64 pipe"]
    nvar = w.nvar
    # the step against HBM on the algorithmic bytes of the workload, B_alg =
    # 2 NVAR 8 bytes per cell-update (read U^n, write U^n+1 once)
    step_bytes = cells * 2 * nvar * 8
    hbm_frac_step = step_bytes / (ms_step * 1e-3) / HBM_PEAK
    step_traffic = None
    if traffic is not None:
        step_traffic = profiled_step_traffic(summary)
    roof = {"bound": "hbm", "kernel": "one level step (every launch of the step)",
            "achieved": step_bytes / (ms_step * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": hbm_frac_step, "traffic": step_traffic,
            "traffic_source": (f"profiles/{summary}" if step_traffic else None),
            "bytes_per_launch": step_bytes, "ms_per_launch": ms_step,
            "alg_bytes_per_cell_update": 2 * nvar * 8}
    fp64 = fp32 = None
    if traffic is not None:
        fl = profiled_step_flops(summary)
        if fl:
            fp64 = {"flop_per_step": fl, "achieved_tflops": fl / (ms_step * 1e-3) / 1e12,
                    "peak_tflops": FP64_PEAK / 1e12,
                    "frac": fl / (ms_step * 1e-3) / FP64_PEAK}
        fl32 = profiled_step_flops(summary, "FP32") if args.precision == "mixed" else None
        if fl32:
            fp32 = {"flop_per_step": fl32,
                    "achieved_tflops": fl32 / (ms_step * 1e-3) / 1e12,
                    "peak_tflops": FP32_PEAK / 1e12,
                    "frac": fl32 / (ms_step * 1e-3) / FP32_PEAK}

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
        nso, build = native_oracle()
        if nso and "oracle_bind" not in sys.modules:
            os.environ["PC_ORACLE_SO"] = nso
        else:
            build = "portable -O3 -mfma (" + build + ")"
        v, s = cpu_baseline(w, fabs0, dt_cpu, args.cpu_steps, threads)
        model, ncpu = host_cpu()
        # the oracle's Godunov step spreads tiles of full x by 8 by 8 cells
        # over the threads (the reference CPU path's MFIter tiling), its MOL
        # step whole boxes
        if args.mol:
            units, what = len(w.boxes), "boxes"
        else:
            units = sum(-(-(hi[1] - lo[1] + 1) // 8) * -(-(hi[2] - lo[2] + 1) // 8)
                        for lo, hi in w.boxes)
            what = f"tiles (full x by 8 by 8) of {len(w.boxes)} boxes"
        used = min(threads, units)
        cpu = {"value": v, "unit": "Mcell-updates/s", "cores": used,
               "kind": "port",
               "sample": f"{args.cpu_steps} full step(s) of the same level "
                         f"({cells} cells) from the GPU run's state and dt, "
                         f"median of 3 timings ({s:.2f} s), OpenMP over "
                         f"{units} {what}, {threads} threads ({used} busy)",
               "host_cpu": model, "host_logical_cpus": ncpu, "build": build}

    out = {
        "metric": METRIC, "value": value, "unit": "Mcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64 (mixed)",
        "data": "synthetic",
        "config": {"workload": desc, "valid_cells": cells,
                   "boxes": len(w.boxes), "nspec": w.nspec,
                   "parallelism": "single GPU, all boxes per launch",
                   "autotuned": bool(args.autotune), "graph_steps": args.graph,
                   "branch_streams": bool(args.branches),
                   "plans": plans},
        "roofline": roof,
        "roofline_dominant_kernel": roof_kernel,
        "hbm_frac_step": hbm_frac_step,
        "fp64_step": fp64,
        "fp32_step": fp32,
        "kernel_ms": {k: round(v, 4) for k, v in kt.items()},
        "cpu_baseline": cpu,
    }
    lev.close()
    if args.error_steps:
        out["precision_check"] = precision_check(
            ctx, w, sorted(int(x) for x in args.error_steps.split(",")),
            not args.no_cpu_baseline)
    default_line = (args.workload == "premixed" and n == 128 and args.box == 32
                    and not args.mol and args.precision == "fp64")
    if default_line and not args.no_forms:
        out["forms"] = forms(args, ctx, w, ms_step)
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
