"""pc_hydro_source_box_noalloc_v1 (workspace in the library's static device
array, no device allocation) against pc_hydro_source_box_v1 (context
workspace) on the same inputs, through the C ABI: fluxes and hydro source
bitwise equal, on two calls in a row, the second on a smaller box; a box
whose workspace exceeds the static array is refused with a status code."""
import ctypes as C

import numpy as np
import pytest

from particleclem_amd import _lib
from particleclem_amd import workloads as wl
from particleclem_amd.ctypes_types import Array4, make_box, make_params, null_a4
from particleclem_amd.hydro import HydroContext, a4_from_tensor

pytestmark = pytest.mark.gpu


def _args(kind, n, src_scale, seed):
    import torch
    w = wl.sod(n=n, nsteps=0) if kind == 1 else wl.premixed(n=n, box=n, nsteps=0)
    import oracle_bind as ob
    H = ob.HostLevel(w.nspec, w.boxes, w.domlo, w.domhi, w.bclo, w.bchi, w.dx)
    H.set_prims(w)
    H.fill_ghosts()
    U = np.ascontiguousarray(H.fab(0))
    nv = U.shape[0]
    lo, hi = w.boxes[0]
    g4 = [x - 4 for x in lo]
    rng = np.random.default_rng(seed)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    src = None
    if src_scale:
        s = src_scale * U * rng.standard_normal(U.shape)
        s[6] = 0.0
        src = T(s)
    dx = w.dx
    dt = 0.2 * dx[0] / (3e5 if kind == 9 else 1.5)
    m = [hi[d] - lo[d] + 1 for d in range(3)]
    shp = lambda d: [m[2] + (d == 2), m[1] + (d == 1), m[0] + (d == 0)]
    av = [dx[1] * dx[2], dx[2] * dx[0], dx[0] * dx[1]]
    area = [T(np.full([1] + shp(d), av[d])) for d in range(3)]
    vol = T(np.full((1, m[2], m[1], m[0]), dx[0] * dx[1] * dx[2]))
    return dict(w=w, U=T(U), src=src, lo=lo, hi=hi, g4=g4, nv=nv, dx=dx, dt=dt,
                area=area, vol=vol, shp=shp)


def _call(ctx, fn, a):
    import torch
    flx = [torch.zeros([a["nv"]] + a["shp"](d), dtype=torch.float64, device="cuda")
           for d in range(3)]
    m = [a["hi"][d] - a["lo"][d] + 1 for d in range(3)]
    hs = torch.full((a["nv"], m[2], m[1], m[0]), 0.25, dtype=torch.float64, device="cuda")
    w = a["w"]
    i3 = lambda v: (C.c_int * 3)(*v)
    prm = make_params(ppm_type=1)
    F = (Array4 * 3)(*[a4_from_tensor(f, a["lo"]) for f in flx])
    A = (Array4 * 3)(*[a4_from_tensor(x, a["lo"]) for x in a["area"]])
    rc = getattr(ctx.L, fn)(
        ctx.h, make_box(a["lo"], a["hi"]), i3(w.domlo), i3(w.domhi), i3(w.bclo),
        i3(w.bchi), a4_from_tensor(a["U"], a["g4"]),
        a4_from_tensor(a["src"], a["g4"]) if a["src"] is not None else null_a4(),
        a4_from_tensor(hs, a["lo"]), (C.c_double * 3)(*a["dx"]), a["dt"], C.byref(prm),
        F, A, a4_from_tensor(a["vol"], a["lo"]), None)
    return rc, flx, hs


def _same(ctx, a, label):
    import torch
    rc0, f0, h0 = _call(ctx, "pc_hydro_source_box_v1", a)
    rc1, f1, h1 = _call(ctx, "pc_hydro_source_box_noalloc_v1", a)
    torch.cuda.synchronize()
    assert rc0 == 0 and rc1 == 0, (rc0, rc1, ctx.L.pc_last_error(ctx.h))
    nd = sum(int((x != y).sum()) for x, y in zip(f0 + [h0], f1 + [h1]))
    print(f"{label}: {nd} values differ between the two entry points")
    assert nd == 0
    assert all(bool(torch.isfinite(x).all()) for x in f1 + [h1])


@pytest.mark.parametrize("kind", [9, 1])
def test_noalloc_equals_context_entry_point(cuda, kind):
    ctx = HydroContext(kind)
    big = _args(kind, 32, 1e-2 if kind == 1 else 2e-4 / 2e-8, seed=3)
    _same(ctx, big, f"eos {kind}, 32^3 box with sources")
    small = _args(kind, 16, 0.0, seed=4)
    _same(ctx, small, f"eos {kind}, 16^3 box after it")
    ctx.close()


def test_noalloc_refuses_boxes_beyond_its_scratch(cuda):
    ctx = HydroContext(9)
    a = _args(9, 40, 0.0, seed=5)
    rc, _, _ = _call(ctx, "pc_hydro_source_box_noalloc_v1", a)
    assert rc == -7, rc
    assert b"allocation-free scratch" in ctx.L.pc_last_error(ctx.h)
    ctx.close()
