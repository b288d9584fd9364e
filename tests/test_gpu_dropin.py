"""The drop-in box drivers through the C ABI against the CPU oracle on the
same inputs: pc_ctoprim_v1 + pc_umdrv_v1 (the reference's pc_umdrv argument
list, caller-provided q/qaux with a FAB larger than grow(bx,4) and a
non-zero uout to accumulate into) and pc_hydro_source_box_v1 (one box of
construct_hydro_source, with and without sources), for wall, symmetry,
outflow, periodic and inflow (low-order boundary) faces.  Fluxes and dU/dt
are compared to 1e-12 of each component's scale, dU/dt by the error it
makes in the updated state U + dt dU/dt (gamma-law: bitwise equal)."""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob
from particleclem_amd import workloads as wl
from particleclem_amd.ctypes_types import Array4, a4_from_numpy, make_box, make_params
from parity_util import compare, rate_scale

pytestmark = pytest.mark.gpu


def host_state(w, bclo, bchi):
    """Valid state of box 0 of workload w with ghosts filled by the oracle's
    level ghost fill under the given BCs."""
    H = ob.HostLevel(w.nspec, w.boxes, w.domlo, w.domhi, bclo, bchi, w.dx)
    H.set_prims(w)
    H.fill_ghosts()
    return np.ascontiguousarray(H.fab(0))


def faces(lo, hi, d):
    h = list(hi)
    h[d] += 1
    return lo, h


def zeros(nc, lo, hi):
    return np.zeros((nc, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1))


def run_box(kind, w, bclo, bchi, src_scale=0.0, ncalls=1, ppm=1):
    import torch
    from particleclem_amd.hydro import HydroContext
    lo, hi = w.boxes[0]
    g4 = ([x - 4 for x in lo], [x + 4 for x in hi])
    U = host_state(w, bclo, bchi)
    nv = U.shape[0]
    rng = np.random.default_rng(1)
    src = None
    if src_scale:
        src = src_scale * U * rng.standard_normal(U.shape)
        src[6] = 0.0
    dx = w.dx
    dt = 0.2 * dx[0] / (3e5 if kind == 9 else 1.5)
    prm = make_params(ppm_type=ppm)
    vol = np.full((1,) + U.shape[1:], dx[0] * dx[1] * dx[2])[:, 4:-4, 4:-4, 4:-4].copy()
    area = []
    for d in range(3):
        fl, fh = faces(lo, hi, d)
        area.append(np.full(zeros(1, fl, fh).shape, [dx[1] * dx[2], dx[2] * dx[0], dx[0] * dx[1]][d]))
    # ----- oracle
    flx_r = [zeros(nv, *faces(lo, hi, d)) for d in range(3)]
    hs_r = zeros(nv, lo, hi) + 0.25
    A4 = lambda a, l: a4_from_numpy(a, l)
    rc = ob.lib().pc_ref_hydro_source_box(
        kind, make_box(lo, hi), ob.iarr(w.domlo), ob.iarr(w.domhi), ob.iarr(bclo),
        ob.iarr(bchi), A4(U, g4[0]), A4(src, g4[0]) if src is not None else Array4(),
        A4(hs_r, lo), ob.darr(dx), dt, C.byref(prm),
        (Array4 * 3)(*[A4(flx_r[d], lo) for d in range(3)]),
        (Array4 * 3)(*[A4(area[d], lo) for d in range(3)]), A4(vol, lo))
    assert rc == 0
    # ----- GPU
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    ctx = HydroContext(kind)
    dU = T(U)
    dsrc = T(src) if src is not None else None
    dflx = [T(zeros(nv, *faces(lo, hi, d))) for d in range(3)]
    dar = [T(a) for a in area]
    dvol = T(vol)
    for _ in range(ncalls):
        dhs = T(zeros(nv, lo, hi) + 0.25)
        ctx.hydro_source_box((lo, hi), w.domlo, w.domhi, bclo, bchi, (dU, g4[0]),
                             (dsrc, g4[0]) if dsrc is not None else None,
                             (dhs, lo), dx, dt, prm,
                             [(f, lo) for f in dflx], [(a, lo) for a in dar],
                             (dvol, lo))
    torch.cuda.synchronize()
    worst = 0.0
    errs = [compare(dflx[d].cpu().numpy(), flx_r[d]) for d in range(3)]
    errs.append(compare(dhs.cpu().numpy(), hs_r, scale=rate_scale(U, dt)))
    print(f"hydro_source_box eos={kind} bc={bclo}/{bchi} src={src_scale}: "
          f"rel err flx x/y/z, dU/dt: {[f'{e:.2e}' for e, _, _ in errs]}; "
          f"{sum(n for _, n, _ in errs)}/{sum(t for _, _, t in errs)} not bitwise")
    ctx.close()
    # FP64: every flux and rate value bit for bit the oracle's
    assert sum(n for _, n, _ in errs) == 0, [n for _, n, _ in errs]
    return max(e for e, _, _ in errs)


@pytest.mark.parametrize("kind", [1, 9])
def test_hydro_source_box_periodic_outflow(cuda, kind):
    w = wl.sod(n=16, nsteps=0) if kind == 1 else wl.premixed(n=16, box=16, nsteps=0)
    worst = run_box(kind, w, w.bclo, w.bchi)
    assert worst <= 1e-12, worst


def test_hydro_source_box_walls_and_inflow(cuda):
    w = wl.premixed(n=16, box=16, nsteps=0)
    # x: symmetry / no-slip wall, y: inflow (low-order boundary) / slip wall,
    # z: outflow
    bclo = [3, 1, 2]
    bchi = [5, 4, 2]
    assert run_box(9, w, bclo, bchi) <= 1e-12
    w1 = wl.sod(n=16, nsteps=0)
    assert run_box(1, w1, [1, 3, 0], [6, 5, 0], ppm=0) <= 1e-12


def test_hydro_source_box_with_sources(cuda):
    w = wl.sod(n=16, nsteps=0)
    assert run_box(1, w, w.bclo, w.bchi, src_scale=1e-2) <= 1e-12
    w9 = wl.premixed(n=16, box=16, nsteps=0)
    assert run_box(9, w9, w9.bclo, w9.bchi, src_scale=2e-4 / 2e-8) <= 1e-12


def test_repeated_calls_reuse_job_slots(cuda):
    w = wl.sod(n=8, nsteps=0)
    assert run_box(1, w, w.bclo, w.bchi, ncalls=70) <= 1e-12


@pytest.mark.parametrize("with_src", [False, True])
def test_umdrv_with_caller_primitives(cuda, with_src):
    """pc_ctoprim_v1 on a q FAB grown by 5, pc_srctoprim_v1 on grow(bx,4)
    (with_src), then pc_umdrv_v1 accumulating into a non-zero uout: matches
    pc_ref_ctoprim + pc_ref_srctoprim + pc_ref_umdrv."""
    import torch
    from particleclem_amd.hydro import HydroContext
    w = wl.premixed(n=16, box=16, nsteps=0)
    kind = 9
    lo, hi = w.boxes[0]
    U = host_state(w, w.bclo, w.bchi)
    nv, qv = U.shape[0], U.shape[0] + 1
    g4 = ([x - 4 for x in lo], [x + 4 for x in hi])
    g5 = ([x - 5 for x in lo], [x + 5 for x in hi])
    U5 = np.zeros((nv,) + tuple(s + 2 for s in U.shape[1:]))
    U5[:, 1:-1, 1:-1, 1:-1] = U
    U5[:, 0], U5[:, -1] = U5[:, 1], U5[:, -2]
    U5[:, :, 0], U5[:, :, -1] = U5[:, :, 1], U5[:, :, -2]
    U5[:, :, :, 0], U5[:, :, :, -1] = U5[:, :, :, 1], U5[:, :, :, -2]
    dx = w.dx
    dt = 2e-8
    A4 = a4_from_numpy
    q_r = zeros(qv, *g5)
    qa_r = zeros(6, *g5)
    assert ob.lib().pc_ref_ctoprim(kind, make_box(*g5), A4(U5, g5[0]), A4(q_r, g5[0]),
                                   A4(qa_r, g5[0])) == 0
    src = srcq_r = None
    if with_src:
        # a source that changes the state by ~1e-4 of itself over dt/2
        src = np.ascontiguousarray(
            2e-4 / dt * U5 * np.random.default_rng(7).standard_normal(U5.shape))
        srcq_r = zeros(qv, *g5)
        assert ob.lib().pc_ref_srctoprim(kind, make_box(*g4), A4(q_r, g5[0]),
                                         A4(qa_r, g5[0]), A4(src, g5[0]),
                                         A4(srcq_r, g5[0])) == 0
    flx_r = [zeros(nv, *faces(lo, hi, d)) for d in range(3)]
    uout_r = zeros(nv, lo, hi) + 1.5
    area = [np.full(zeros(1, *faces(lo, hi, d)).shape, dx[0] * dx[1]) for d in range(3)]
    vol = np.full(zeros(1, lo, hi).shape, dx[0] ** 3)
    assert ob.lib().pc_ref_umdrv(
        kind, 0.0, make_box(lo, hi), ob.iarr(w.domlo), ob.iarr(w.domhi),
        ob.iarr(w.bclo), ob.iarr(w.bchi), A4(U5, g5[0]), A4(uout_r, lo),
        A4(q_r, g5[0]), A4(qa_r, g5[0]),
        A4(srcq_r, g5[0]) if with_src else Array4(), ob.darr(dx), dt, 1, 4, 1, 0,
        1, 0.1, (Array4 * 3)(*[A4(f, lo) for f in flx_r]),
        (Array4 * 3)(*[A4(a, lo) for a in area]), A4(vol, lo), None) == 0
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    ctx = HydroContext(kind)
    dU5 = T(U5)
    dq = T(zeros(qv, *g5))
    dqa = T(zeros(6, *g5))
    ctx.ctoprim(g5, (dU5, g5[0]), (dq, g5[0]), (dqa, g5[0]))
    dsrcq = None
    if with_src:
        dsrcq = T(zeros(qv, *g5))
        ctx.srctoprim(g4, (dq, g5[0]), (dqa, g5[0]), (T(src), g5[0]),
                      (dsrcq, g5[0]))
    dflx = [T(zeros(nv, *faces(lo, hi, d))) for d in range(3)]
    duout = T(zeros(nv, lo, hi) + 1.5)
    ctx.umdrv((lo, hi), w.domlo, w.domhi, w.bclo, w.bchi, (dU5, g5[0]),
              (duout, lo), (dq, g5[0]), (dqa, g5[0]),
              (dsrcq, g5[0]) if with_src else None, dx, dt, 1, 4, 1, 0.1,
              [(f, lo) for f in dflx], [(T(a), lo) for a in area], (T(vol), lo))
    torch.cuda.synchronize()
    c = [compare(dq.cpu().numpy(), q_r), compare(dqa.cpu().numpy(), qa_r),
         compare(duout.cpu().numpy(), uout_r, scale=rate_scale(U, dt))]
    if with_src:
        c.append(compare(dsrcq.cpu().numpy(), srcq_r))
    c += [compare(dflx[d].cpu().numpy(), flx_r[d]) for d in range(3)]
    e = [x[0] for x in c]
    print("umdrv drop-in errors", e, "not bitwise", [x[1] for x in c])
    assert max(e) <= 1e-12, e
    assert all(x[1] == 0 for x in c), [x[1] for x in c]
    ctx.close()


def test_box_calls_rotated_over_streams(cuda):
    """pc_hydro_source_box_v1 for the boxes of a level issued on four
    streams in rotation (as AMReX's MFIter does on GPUs; the calls of
    different streams overlap) gives what the same calls give on one stream,
    bitwise: each stream has its own workspace in the context."""
    import torch
    from particleclem_amd.hydro import DeviceLevel, HydroContext
    w = wl.premixed(n=32, box=16, nsteps=0)
    ctx = HydroContext(9)
    lev = DeviceLevel.from_workload(ctx, w)
    lev.init_dt()
    lev.fill_ghosts()
    dt = lev.get_dt()
    nv = w.nvar
    dx = w.dx
    prm = make_params(**w.params)
    U = []
    for b in range(len(w.boxes)):
        u = torch.empty(lev.fab_shape(b), dtype=torch.float64, device="cuda")
        u.copy_(torch.from_numpy(lev.read_box(b)))
        U.append(u)
    lev.close()
    streams = [torch.cuda.Stream() for _ in range(4)]

    def run(rotate, reps):
        outs = []
        for b, (lo, hi) in enumerate(w.boxes):
            shp = tuple(hi[d] - lo[d] + 1 for d in (2, 1, 0))
            hs = torch.zeros((nv,) + shp, dtype=torch.float64, device="cuda")
            fl, ar = [], []
            for d in range(3):
                fs = list(shp)
                fs[2 - d] += 1
                fl.append(torch.zeros((nv,) + tuple(fs), dtype=torch.float64, device="cuda"))
                ar.append(torch.full((1,) + tuple(fs), dx[0] * dx[0], dtype=torch.float64,
                                     device="cuda"))
            vol = torch.full((1,) + shp, dx[0] ** 3, dtype=torch.float64, device="cuda")
            outs.append((lo, hi, hs, fl, ar, vol))
        torch.cuda.synchronize()
        for _ in range(reps):
            for b, (lo, hi, hs, fl, ar, vol) in enumerate(outs):
                s = streams[b % 4] if rotate else None
                ctx.hydro_source_box((lo, hi), w.domlo, w.domhi, w.bclo, w.bchi,
                                     (U[b], [x - 4 for x in lo]), None, (hs, lo), dx, dt,
                                     prm, [(f, lo) for f in fl], [(a, lo) for a in ar],
                                     (vol, lo), stream=s)
        torch.cuda.synchronize()
        return [[o[2].cpu().numpy()] + [f.cpu().numpy() for f in o[3]] for o in outs]

    # the hydro source accumulates (uout += dU/dt): both runs make 3 rounds
    one = run(False, 3)
    many = run(True, 3)
    ndiff = sum(int(np.count_nonzero(a != b)) for x, y in zip(one, many)
                for a, b in zip(x, y))
    print(f"8 boxes over 4 streams x 3 rounds: {ndiff} values differ from one stream")
    assert ndiff == 0
    ctx.close()
