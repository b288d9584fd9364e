"""The mixed-precision level step (FP32 reconstruction, limiters, Riemann
solves and intermediates; FP64 EOS inversion, internal energy, flux assembly
and update) against the FP64 step, which is bitwise the oracle.  Tolerance
(relative L1 over the valid cells): 1e-6 in rho and T, the bound the
mixed-precision form is built to."""
import numpy as np
import pytest

from particleclem_amd import workloads as wl
from particleclem_amd.precision import error_norms

pytestmark = pytest.mark.gpu


def level(w, mixed, **over):
    from particleclem_amd.hydro import DeviceLevel, HydroContext
    ctx = HydroContext(w.nspec)
    lev = DeviceLevel.from_workload(ctx, w, **over)
    lev.set_precision(mixed)
    return ctx, lev


@pytest.mark.parametrize("ppm_type", [1, 0])
def test_mixed_premixed_close_to_fp64(cuda, ppm_type):
    w = wl.premixed(n=32, box=16, nsteps=0)
    over = dict(ppm_type=ppm_type, fixed_dt=2.0e-8)
    c64, l64 = level(w, False, **over)
    c32, l32 = level(w, True, **over)
    for x in (l64, l32):
        x.init_dt()
        x.step(20)
    e = error_norms(l32, l64, w.nspec, len(w.boxes))
    print(f"ppm={ppm_type} 20 steps: rho L1 {e['rho']['L1']:.2e} Linf "
          f"{e['rho']['Linf']:.2e}; T L1 {e['T']['L1']:.2e} Linf {e['T']['Linf']:.2e}; "
          f"max Y L1 {max(e[f'Y{k}']['L1'] for k in range(9)):.2e}")
    assert e["rho"]["L1"] <= 1e-6 and e["T"]["L1"] <= 1e-6
    assert l32.status()[0] == 0
    for x in (l64, l32):
        x.close()
    c64.close()
    c32.close()


def test_mixed_conserves_on_periodic_box(cuda):
    """Fully periodic Sod: sum of rho, rho u, rho E is kept to FP64 round-off
    in the mixed step (the FP64 update telescopes the same face fluxes)."""
    import torch
    w = wl.sod(n=32, nsteps=0, xbc="interior")
    ctx, lev = level(w, True)
    lev.init_dt()

    def totals():
        t = torch.zeros(5, dtype=torch.float64, device="cuda")
        for b in range(len(w.boxes)):
            U = lev.valid_tensor(b)
            t += U[:5].reshape(5, -1).sum(dim=1)
        return t.cpu().numpy()

    t0 = totals()
    lev.step(10)
    t1 = totals()
    rel = np.abs(t1 - t0) / np.maximum(np.abs(t0), 1.0)
    print("mass/momentum/energy drift", rel)
    assert rel[0] <= 1e-13 and rel[4] <= 1e-13
    assert np.all(rel[1:4] <= 1e-12)
    lev.close()
    ctx.close()


def test_mixed_ragged_with_graph(cuda):
    """The mixed step through the batched path (ragged table, hipGraph)
    equals the mixed step launched directly, bitwise."""
    w = wl.ragged(nboxes=8, sides=(8, 16, 24), probs=(0.4, 0.4, 0.2))
    ca, a = level(w, True)
    cb, b = level(w, True)
    for x in (a, b):
        x.init_dt()
    b.set_graph(2)
    a.step(5)
    b.step(5)
    for k in range(len(w.boxes)):
        assert np.array_equal(a.read_box(k), b.read_box(k))
    for x in (a, b):
        x.close()
    ca.close()
    cb.close()


def test_precision_switch_replans_transverse_stage(cuda):
    """The transverse stage runs two threads per face in FP64 and one in
    FP32, so switching a level's precision re-plans it, keeping a shape set
    before the switch: a level with tiles for that stage, switched FP64 ->
    mixed -> FP64, equals a level with the default plan switched the same
    way, bitwise after each phase, and its plan has twice the blocks in
    FP64."""
    w = wl.premixed(n=32, box=16, nsteps=0)
    over = dict(fixed_dt=2.0e-8)
    c0, ref64 = level(w, False, **over)
    c1, ref32 = level(w, False, **over)
    c2, sw = level(w, False, **over)
    sw.set_plan("trans_xy", (5, 3, 0), 0)
    for x in (ref64, ref32, sw):
        x.init_dt()
    ref64.step(2)
    ref32.step(2)
    sw.step(2)
    assert all(np.array_equal(sw.read_box(b), ref64.read_box(b)) for b in range(len(w.boxes)))
    # both switch to mixed for two steps, then back (ref32 keeps runs of
    # 256 faces for its transverse stage)
    ref32.set_precision(True)
    sw.set_precision(True)
    ref32.step(2)
    sw.step(2)
    assert all(np.array_equal(sw.read_box(b), ref32.read_box(b)) for b in range(len(w.boxes)))
    n32 = sw.plans()["trans_xy"][2]
    sw.set_precision(False)
    ref32.set_precision(False)
    sw.step(2)
    ref32.step(2)
    assert all(np.array_equal(sw.read_box(b), ref32.read_box(b)) for b in range(len(w.boxes)))
    shape, _, n64, _ = sw.plans()["trans_xy"]
    assert shape == (32, 8, 1) and n64 == 2 * n32, (shape, n64, n32)
    assert sw.status()[0] == 0
    for x in (ref64, ref32, sw):
        x.close()
    for c in (c0, c1, c2):
        c.close()
