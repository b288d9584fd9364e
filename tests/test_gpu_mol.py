"""The method-of-lines path on the GPU (through the C ABI) against the CPU
oracle on the same seeded inputs: whole MOL level steps (do_mol_advance,
both Heun stages, ghost values, computeTemp, dt) for the gamma-law gas and
the 9-species mixture, with the approximate Riemann solver and with the
Lax-Friedrichs flux, with and without slopes; the one-box MOL source term
pc_mol_src_box_v1; launch shapes, autotuning and hipGraph replay of MOL
steps (bitwise against direct steps); and the mixed-precision MOL step
against the FP64 one.  Bar: 1e-12 of each component's scale (the FP64
kernels reproduce the oracle's roundings, so values are bitwise equal)."""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob
from particleclem_amd import workloads as wl
from particleclem_amd.ctypes_types import make_params
from parity_util import VALID, assert_parity, compare, rate_scale

pytestmark = pytest.mark.gpu

MOL = dict(do_mol=1)
LAXF = dict(do_mol=1, use_laxf_flux=1)


@pytest.mark.parametrize("over", [MOL, LAXF, dict(do_mol=1, plm_iorder=1)],
                         ids=["riemann", "laxf", "no-slopes"])
def test_sod_mol_matches_oracle(cuda, over):
    w = wl.sod(n=16, nsteps=3)
    assert_parity(w, 3, over, label=f"sod16 {over}")


def test_sod_mol_multibox_slipwall(cuda):
    """The sod-3 configuration (MOL, CFL 0.3, slip walls at the tube ends)
    on several boxes: ghost exchange between boxes in both stages."""
    w = wl.sod(n=32, ny=16, nz=16, max_grid=16, nsteps=2, xbc="slipwall")
    assert len(w.boxes) > 1
    assert_parity(w, 2, MOL, label="sod32 slipwall multibox")


@pytest.mark.parametrize("over", [MOL, LAXF], ids=["riemann", "laxf"])
def test_premixed_mol_9species(cuda, over):
    w = wl.premixed(n=32, box=16, nsteps=2)
    assert_parity(w, 2, over, label=f"premixed 32^3 MOL {over}")


def test_ragged_mol_matches_oracle(cuda):
    w = wl.ragged(nboxes=10, sides=(8, 16, 24), probs=(0.4, 0.4, 0.2))
    assert_parity(w, 2, MOL, label="ragged MOL")


@pytest.mark.parametrize("over", [MOL, LAXF], ids=["riemann", "laxf"])
def test_mol_odd_boxes_every_boundary_kind(cuda, over):
    """Boxes of odd, unequal sides with inflow, no-slip, slip, symmetry and
    outflow faces (the ghost fill's first-order extrapolation and
    reflections in both Heun stages)."""
    w = wl.sod(n=21, ny=13, nz=7, max_grid=11, nsteps=2, xbc="slipwall")
    w.bclo = [1, 5, 3]
    w.bchi = [2, 4, 3]
    assert len(w.boxes) > 2
    assert_parity(w, 3, over, label=f"sod 21x13x7 MOL, all BC kinds {over}")


@pytest.mark.parametrize("over", [MOL, LAXF], ids=["riemann", "laxf"])
@pytest.mark.parametrize("idir", [0, 1])
def test_mol_large_odd_box_partial_tiles(cuda, over, idir):
    """One box of 45 x 37 x 19 cells (the tube along x, or 19 x 45 x 37 with
    it along y): x-face rows longer than a wave and fewer than a workgroup's
    worth per tile, y- and z-face tiles split along x (32 + the rest) and
    along the direction with a shorter last tile, which the FP64 face tiles
    staged in LDS must cover exactly."""
    w = wl.sod(n=45, ny=37, nz=19, nsteps=2, idir=idir)
    assert len(w.boxes) == 1
    assert_parity(w, 2, over, label=f"sod one odd box idir={idir} MOL {over}")


def _box_inputs(kind):
    w = wl.sod(n=16, nsteps=0) if kind == 1 else wl.premixed(n=16, box=16, nsteps=0)
    H = ob.HostLevel(w.nspec, w.boxes, w.domlo, w.domhi, w.bclo, w.bchi, w.dx)
    H.set_prims(w)
    H.fill_ghosts()
    return w, np.ascontiguousarray(H.fab(0))


@pytest.mark.parametrize("kind", [1, 9])
@pytest.mark.parametrize("laxf", [0, 1])
def test_mol_src_box_matches_oracle(cuda, kind, laxf):
    import torch
    from particleclem_amd.hydro import HydroContext
    w, U = _box_inputs(kind)
    lo, hi = w.boxes[0]
    g4 = [x - 4 for x in lo]
    prm = make_params(do_mol=1, use_laxf_flux=laxf)
    src_r, flx_r = ob.mol_src_box(kind, U, lo, hi, w.dx, prm, want_flux=True)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    nv = U.shape[0]
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    dx = w.dx
    area_v = [dx[1] * dx[2], dx[2] * dx[0], dx[0] * dx[1]]
    areas, flx = [], []
    for d in range(3):
        m = list(n)
        m[d] += 1
        areas.append(T(np.full((1, m[2], m[1], m[0]), area_v[d])))
        flx.append(T(np.full((nv, m[2], m[1], m[0]), np.nan)))
    vol = T(np.full((1, n[2], n[1], n[0]), dx[0] * dx[1] * dx[2]))
    ctx = HydroContext(kind)
    dU = T(U)
    out = T(np.full((nv, n[2], n[1], n[0]), np.nan))
    ctx.mol_src_box((lo, hi), (dU, g4), (out, lo), prm, [(a, lo) for a in areas],
                    (vol, lo), flx=[(f, lo) for f in flx])
    out2 = T(np.full((nv, n[2], n[1], n[0]), np.nan))
    ctx.mol_src_box((lo, hi), (dU, g4), (out2, lo), prm, [(a, lo) for a in areas],
                    (vol, lo))
    torch.cuda.synchronize()
    dt = 0.3 * dx[0] / (3e5 if kind == 9 else 1.5)
    e_src = compare(out.cpu().numpy(), src_r, scale=rate_scale(U[:, 4:-4, 4:-4, 4:-4], dt))
    e_flx = [compare(flx[d].cpu().numpy(), flx_r[d]) for d in range(3)]
    print(f"mol_src_box eos={kind} laxf={laxf}: src {e_src[0]:.2e} ({e_src[1]} not "
          f"bitwise), flx {[f'{e:.2e}' for e, _, _ in e_flx]}")
    assert e_src[0] <= 1e-12
    assert all(e <= 1e-12 for e, _, _ in e_flx)
    assert e_src[1] == 0 and all(n == 0 for _, n, _ in e_flx)
    assert np.array_equal(out2.cpu().numpy(), out.cpu().numpy())
    ctx.close()


def _level(w, **over):
    from particleclem_amd.hydro import DeviceLevel, HydroContext
    ctx = HydroContext(w.nspec)
    lev = DeviceLevel.from_workload(ctx, w, **over)
    lev.init_dt()
    return ctx, lev


def _states(lev, nb):
    return [lev.read_box(b) for b in range(nb)]


def test_mol_launch_shapes_autotune_and_graph_are_bitwise(cuda):
    from particleclem_amd import _lib
    w = wl.ragged(nboxes=8, sides=(8, 16, 24), probs=(0.4, 0.4, 0.2))
    nb = len(w.boxes)
    c0, ref = _level(w, **MOL)
    ref.step(3)
    want = _states(ref, nb)
    mol_stages = [s for s in _lib.STAGES if s.startswith("mol_") or s == "ghost"]
    for shape, persist in [((5, 2, 1), 0), ((4, 4, 0), 2), (None, 4)]:
        c, lev = _level(w, **MOL)
        for st in mol_stages:
            lev.set_plan(st, shape, persist)
        lev.step(3)
        got = _states(lev, nb)
        assert all(np.array_equal(a, b) for a, b in zip(got, want)), (shape, persist)
        assert lev.get_dt() == ref.get_dt()
        lev.close()
        c.close()
    c, lev = _level(w, **MOL)
    before = _states(lev, nb)
    lev.autotune(reps=1)
    assert all(np.array_equal(a, b) for a, b in zip(_states(lev, nb), before))
    lev.set_graph(2)
    lev.step(3)
    got = _states(lev, nb)
    assert all(np.array_equal(a, b) for a, b in zip(got, want))
    assert lev.get_dt() == ref.get_dt()
    assert lev.status()[0] == 0
    lev.close()
    c.close()
    ref.close()
    c0.close()


def test_mol_kernel_times_cover_both_stages(cuda):
    from particleclem_amd import _lib
    w = wl.sod(n=32, ny=16, nz=16, max_grid=16, nsteps=0)
    c, lev = _level(w, **MOL)
    lev.enable_timing(True)
    lev.step(2)
    t = lev.kernel_times()
    assert list(t) == _lib.KERNELS
    used = {k for k, v in t.items() if v > 0}
    print({k: round(v, 4) for k, v in t.items() if v > 0})
    assert {"ctoprim", "mol_x", "mol_y", "mol_z", "mol_update1",
            "mol_update2", "dt"} <= used
    # the primitives read ghost values at their source cells: no fill launch
    assert not used & {"ghost", "divu", "trace_x", "flux1_x", "trans_xy", "final_x",
                       "update"}
    lev.close()
    c.close()


def test_mixed_precision_mol_close_to_fp64(cuda):
    from particleclem_amd.precision import error_norms
    w = wl.premixed(n=32, box=16, nsteps=0)
    over = dict(do_mol=1, fixed_dt=2.0e-8)
    c64, l64 = _level(w, **over)
    c32, l32 = _level(w, **over)
    l32.set_precision(True)
    for x in (l64, l32):
        x.step(20)
    e = error_norms(l32, l64, w.nspec, len(w.boxes))
    print(f"MOL mixed 20 steps: rho L1 {e['rho']['L1']:.2e} T L1 {e['T']['L1']:.2e}")
    assert e["rho"]["L1"] <= 1e-6 and e["T"]["L1"] <= 1e-6
    for x in (l64, l32):
        x.close()
    c64.close()
    c32.close()


def test_mol_precision_switch_replans_face_fluxes(cuda):
    """The FP64 MOL fluxes run on face tiles staged in LDS, the mixed y and z
    fluxes one thread per face, so switching a MOL level's precision
    re-plans those stages: a level switched FP64 -> mixed -> FP64 between
    steps equals a level stepped the same way with the plans made fresh at
    each switch (a new level from the state), bitwise after each phase."""
    w = wl.premixed(n=32, box=16, nsteps=0)
    over = dict(do_mol=1, fixed_dt=2.0e-8)
    c, sw = _level(w, **over)
    assert sw.plans()["mol_y"][0] == "tiles" and sw.plans()["mol_x"][0] == "tiles"
    sw.step(2)
    sw.set_precision(True)
    assert sw.plans()["mol_y"][0] == "runs" and sw.plans()["mol_x"][0] == "tiles"
    sw.step(2)
    mid = _states(sw, len(w.boxes))
    sw.set_precision(False)
    assert sw.plans()["mol_y"][0] == "tiles"
    sw.step(2)
    got = _states(sw, len(w.boxes))
    # the same phases on fresh levels loaded with the state at each switch
    c1, a = _level(w, **over)
    a.step(2)
    s2 = _states(a, len(w.boxes))
    c2, b = _level(w, **over)
    b.set_precision(True)
    for k, f in enumerate(s2):
        b.write_box(k, f)
    b.step(2)
    assert all(np.array_equal(x, y) for x, y in zip(_states(b, len(w.boxes)), mid))
    c3, d = _level(w, **over)
    for k, f in enumerate(mid):
        d.write_box(k, f)
    d.step(2)
    assert all(np.array_equal(x, y) for x, y in zip(_states(d, len(w.boxes)), got))
    assert sw.status()[0] == 0
    for x, cc in ((sw, c), (a, c1), (b, c2), (d, c3)):
        x.close()
        cc.close()


def test_mol_option_errors(cuda):
    from particleclem_amd._lib import HydroError
    from particleclem_amd.hydro import DeviceLevel, HydroContext
    w = wl.sod(n=8, nsteps=0)
    ctx = HydroContext(1)
    with pytest.raises(HydroError) as e:
        DeviceLevel.from_workload(ctx, w, do_mol=1, mol_iorder=3)
    assert e.value.code == -10
    # do_mol skips the Godunov option checks (ppm_type is not read)
    lev = DeviceLevel.from_workload(ctx, w, do_mol=1, ppm_type=7)
    lev.close()
    ctx.close()
