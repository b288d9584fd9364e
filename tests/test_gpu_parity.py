"""GPU path (through the C ABI) against the CPU oracle on the same seeded
inputs: per-cell state and dt after whole Godunov steps, at sizes the oracle
finishes in seconds.  The bar is bitwise equality of every valid value, dt
and level time (the FP64 kernels reproduce the oracle's roundings)."""
import pytest

from particleclem_amd import workloads as wl
from parity_util import assert_parity

pytestmark = pytest.mark.gpu


def test_sod_ppm_single_box(cuda):
    w = wl.sod(n=16, nsteps=3)
    assert_parity(w, 3, label="sod16 ppm")


def test_sod_ppm_multibox(cuda):
    # 32 x 16 x 16 tube as 4 boxes of 16 x 8 x 16 style split
    w = wl.sod(n=32, ny=16, nz=16, max_grid=16, nsteps=2)
    assert len(w.boxes) > 1
    assert_parity(w, 2, label="sod32 multibox")


def test_sod_plm(cuda):
    w = wl.sod(n=16, nsteps=2)
    assert_parity(w, 2, dict(ppm_type=0), label="sod16 plm")


def test_premixed_9species(cuda):
    w = wl.premixed(n=32, box=16, nsteps=2)
    assert_parity(w, 2, label="premixed 32^3 / 16^3 boxes")


@pytest.mark.parametrize("over,nsteps", [({}, 10), (dict(do_mol=1), 5)],
                         ids=["ctu-10-steps", "mol-5-steps"])
def test_premixed_32cubed_boxes_many_steps(cuda, over, nsteps):
    """The premixed level in 32^3 boxes (8 of them, 64^3 cells) over several
    steps: every cell of the state and every dt of the chain (each step's
    dt comes from the previous step's state on the device)."""
    w = wl.premixed(n=64, box=32, nsteps=nsteps)
    assert len(w.boxes) == 8
    assert_parity(w, nsteps, over, label=f"premixed 64^3 / 32^3 boxes {over}, {nsteps} steps")


@pytest.mark.parametrize("over", [dict(ppm_type=0, plm_iorder=1),
                                  dict(ppm_type=0, plm_iorder=2),
                                  dict(ppm_type=1, use_flattening=0),
                                  dict(ppm_type=0, use_flattening=0),
                                  dict(ppm_type=1, difmag=0.0)],
                         ids=["plm1", "plm2", "ppm-noflat", "plm-noflat", "nodifmag"])
def test_option_variants_premixed(cuda, over):
    """The reference's other hydro options on the 9-species level: PLM
    orders 1 and 2, no flattening, no artificial viscosity."""
    w = wl.premixed(n=16, box=8, nsteps=2)
    assert_parity(w, 2, over, label=f"premixed 16^3 {over}")


@pytest.mark.parametrize("ppm_type", [1, 0], ids=["ppm", "plm"])
def test_premixed_every_boundary_kind(cuda, ppm_type):
    """The 9-species level with an inflow face (the low-order boundary
    fluxes), no-slip, slip and symmetry walls and an outflow face."""
    w = wl.premixed(n=16, box=8, nsteps=2)
    w.bclo = [1, 5, 3]
    w.bchi = [2, 4, 2]
    assert_parity(w, 2, dict(ppm_type=ppm_type), label=f"premixed 16^3, all BC kinds, ppm={ppm_type}")


def test_odd_boxes_walls_both_ends(cuda):
    """Boxes of odd, unequal sides on a domain with slip and no-slip walls
    (reflecting ghost fill and the wall Riemann faces)."""
    w = wl.sod(n=21, ny=13, nz=7, max_grid=11, nsteps=2, xbc="slipwall")
    w.bclo = [4, 5, 3]
    w.bchi = [5, 4, 3]
    assert len(w.boxes) > 2
    assert_parity(w, 3, label="sod 21x13x7, walls")


@pytest.mark.parametrize("over", [dict(do_mol=0), dict(do_mol=1)], ids=["ctu", "mol"])
def test_stop_time_last_step(cuda, over):
    """computeNewDt's stop_time limit on the device: the step that would pass
    stop_time ends on it, later steps take dt = 0; dt chain and level time
    bitwise those of the oracle."""
    w = wl.sod(n=16, nsteps=0)
    from parity_util import oracle_level
    H, prm = oracle_level(w)
    dt0 = H.init_dt(prm)
    over = dict(over, stop_time=2.6 * dt0)
    assert_parity(w, 5, over, label=f"sod16 stop_time {over}")


@pytest.mark.parametrize("idir", [0, 1, 2], ids=["x", "y", "z"])
def test_long_box_unfused_ppm_trace(cuda, idir):
    """A box longer than one workgroup's pencil (300 cells along the tube):
    the PPM trace runs on its LDS tiles with the first fluxes and the cell
    corrections in kernels of their own instead of the fused pencil
    kernel."""
    w = wl.sod(n=300, ny=8, nz=8, idir=idir, nsteps=2)
    assert len(w.boxes) == 1
    assert_parity(w, 2, label=f"sod 300x8x8 along {'xyz'[idir]}")


@pytest.mark.parametrize("over", [{}, dict(do_mol=1)], ids=["ctu", "mol"])
def test_sod_thin_periodic_domain(cuda, over):
    """A tube 2 cells thick in y and 3 in z, periodic there, split along x
    into boxes: ghost cells 3 and 4 cells out lie more than one period away,
    so their values come from the cell that several wraps reach (the
    primitives' ghost-source mapping on the device, the ghost fill in the
    oracle)."""
    w = wl.sod(n=24, ny=2, nz=3, max_grid=8, nsteps=3)
    assert len(w.boxes) == 3
    assert_parity(w, 3, over, label=f"sod 24x2x3 periodic y/z {over}")
