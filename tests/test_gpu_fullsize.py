"""The GPU path at the workloads' full sizes.

* The Sod tube of 64^3 cells as one box, 10 PPM steps, and the premixed
  H2/air level of 128^3 cells as 64 boxes of 32^3 over 20 Godunov steps and
  one method-of-lines step: every value bitwise the oracle's.
* The ragged level of 300 boxes (sides 8-64), each box its own periodic
  domain: the first step against the oracle on one box of each side (the
  boxes are independent domains; the oracle steps them with the level's dt,
  a CPU step of all 300 would take a minute), then 100 steps replayed from
  hipGraphs checked by what holds at any size: mass, momentum and total
  energy of every box are conserved to round-off (the update telescopes the
  face fluxes), the species densities sum to the density in every cell (the
  species fluxes are normalised to the mass flux), the state stays finite.
* The 256^3 level (512 boxes of 32^3) in mixed precision against FP64 after
  1, 10 and 100 steps: relative L1 error in rho and T <= 1e-6.
"""
import numpy as np
import pytest

from particleclem_amd import workloads as wl
from parity_util import VALID, assert_parity, compare, run_oracle

pytestmark = pytest.mark.gpu


def test_sod_64cubed_10_steps(cuda):
    w = wl.sod(n=64, nsteps=10)
    assert len(w.boxes) == 1
    assert_parity(w, 10, label="sod 64^3, 10 PPM steps")


def test_premixed_128cubed_20_steps(cuda):
    """The benchmark's level over 20 Godunov steps: every value of the state
    and every dt of the chain bitwise the oracle's (the oracle takes ~35 s on
    16 host threads)."""
    w = wl.premixed(n=128, box=32, nsteps=20)
    assert len(w.boxes) == 64
    assert_parity(w, 20, label="premixed 128^3 / 64 boxes of 32^3, 20 steps")


def test_premixed_128cubed_mol_first_step(cuda):
    """The method-of-lines step (do_mol_advance, both Heun stages) on the
    same level."""
    w = wl.premixed(n=128, box=32, nsteps=1)
    assert_parity(w, 1, dict(do_mol=1), label="premixed 128^3 MOL")


def _box_totals(lev, nboxes):
    import torch
    t = [lev.valid_tensor(b)[:5].reshape(5, -1).sum(dim=1) for b in range(nboxes)]
    return torch.stack(t).cpu().numpy()


def _species_defect(lev, nboxes, nspec):
    """max over cells of |sum_k rho Y_k - rho| / rho"""
    import torch
    worst = torch.zeros((), dtype=torch.float64, device="cuda")
    for b in range(nboxes):
        U = lev.valid_tensor(b)
        d = (U[7:7 + nspec].sum(dim=0) - U[0]).abs() / U[0]
        worst = torch.maximum(worst, d.max())
    return float(worst)


def test_ragged_300_boxes_oracle_then_invariants(cuda):
    from particleclem_amd.hydro import DeviceLevel, HydroContext
    w = wl.ragged(nboxes=300, nsteps=100)
    nb = len(w.boxes)
    assert nb == 300
    ctx = HydroContext(w.nspec)
    lev = DeviceLevel.from_workload(ctx, w)
    lev.init_dt()
    dt0 = lev.get_dt()
    lev.step(1)
    sides = [w.boxes[b][1][0] + 1 for b in range(nb)]
    pick = [sides.index(s) for s in sorted(set(sides))]
    assert len(pick) == 6
    sub = wl.Workload(w.name, w.nspec, [w.boxes[b] for b in pick], w.domlo, w.domhi,
                      w.bclo, w.bchi, w.dx, w.prim_mode, [w.prims[b] for b in pick],
                      periodic_self=True, params=w.params, nsteps=1)
    ref, _ = run_oracle(sub, 1, dt0=dt0)
    worst, nbits, ntot = 0.0, 0, 0
    for r, b in enumerate(pick):
        e, n, t = compare(lev.read_box(b)[VALID], ref[r][VALID])
        worst, nbits, ntot = max(worst, e), nbits + n, ntot + t
    print(f"[ragged 300 boxes] first step, boxes {pick} (sides "
          f"{[sides[b] for b in pick]}): max rel err {worst:.3e}, "
          f"{nbits} of {ntot} values not bitwise")
    assert worst <= 1e-12
    assert nbits == 0
    del ref

    t0 = _box_totals(lev, nb)
    lev.set_graph(10)
    lev.step(100)
    rc, bad, small = lev.status()
    assert rc == 0 and not bad and not small, (rc, bad, small)
    t1 = _box_totals(lev, nb)
    mass, energy = np.abs(t0[:, 0]), np.abs(t0[:, 4])
    mom = np.sqrt(mass * energy)
    scale = np.stack([mass, mom, mom, mom, energy], axis=1)
    drift = np.max(np.abs(t1 - t0) / scale)
    spec = _species_defect(lev, nb, w.nspec)
    print(f"[ragged 300 boxes] 100 graph steps: max conservation drift {drift:.2e}, "
          f"max |sum rhoY - rho|/rho {spec:.2e}")
    assert drift <= 1e-13
    assert spec <= 1e-12
    lev.close()
    ctx.close()


def test_mixed_256cubed_against_fp64(cuda):
    from particleclem_amd.hydro import DeviceLevel, HydroContext
    from particleclem_amd.precision import error_norms
    w = wl.premixed(n=256, box=32, nsteps=100)
    nb = len(w.boxes)
    assert nb == 512
    ctx = HydroContext(w.nspec)
    levs = []
    for mixed in (False, True):
        lev = DeviceLevel.from_workload(ctx, w)
        lev.set_precision(mixed)
        lev.init_dt()
        levs.append(lev)
    done = 0
    for upto in (1, 10, 100):
        for lev in levs:
            lev.step(upto - done)
        done = upto
        e = error_norms(levs[1], levs[0], w.nspec, nb)
        print(f"[256^3 mixed vs FP64] step {upto}: rho L1 {e['rho']['L1']:.2e} "
              f"Linf {e['rho']['Linf']:.2e}; T L1 {e['T']['L1']:.2e} "
              f"Linf {e['T']['Linf']:.2e}")
        assert e["rho"]["L1"] <= 1e-6 and e["T"]["L1"] <= 1e-6
    for lev in levs:
        assert lev.status()[0] == 0
        lev.close()
    ctx.close()
