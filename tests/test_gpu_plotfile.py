"""pc_level_write_plotfile_v1 on the GPU: the plotfile of a stepped level
read back holds the level's state bitwise and the derived plot variables
equal to the oracle's restatement of Source/Derive.cpp on the same state
(magvort on ghost cells filled by the oracle's fill), for both mechanisms
and a multi-box level; variable order and names are the reference's."""
import numpy as np
import pytest

import oracle_bind as ob
from particleclem_amd import plotfile as pf
from particleclem_amd import workloads as wl
from parity_util import VALID, compare
from test_plotfile import derive

pytestmark = pytest.mark.gpu

SPEC = {1: ["N2"], 9: ["H2", "O2", "H2O", "H", "O", "OH", "HO2", "H2O2", "N2"]}
STATE = ["density", "xmom", "ymom", "zmom", "rho_E", "rho_e", "Temp"]
CODES = {"pressure": (4, 0), "kineng": (3, 0), "soundspeed": (5, 0),
         "MachNumber": (6, 0), "magvort": (7, 0), "logden": (8, 0),
         "x_velocity": (1, 1), "y_velocity": (1, 2), "z_velocity": (1, 3),
         "magvel": (2, 0)}


def all_names(ns):
    sp = SPEC[ns]
    return (STATE + [f"rho_{s}" for s in sp] +
            ["pressure", "kineng", "soundspeed", "MachNumber", "magvort", "logden"] +
            [f"Y({s})" for s in sp] + ["x_velocity", "y_velocity", "z_velocity", "magvel"])


def code_of(name, ns):
    if name in STATE:
        return 0, STATE.index(name)
    if name.startswith("rho_"):
        return 0, 7 + SPEC[ns].index(name[4:])
    if name.startswith("Y("):
        return 1, 7 + SPEC[ns].index(name[2:-1])
    return CODES[name]


@pytest.mark.parametrize("kind", ["premixed", "sod"])
def test_level_plotfile_matches_state_and_oracle(cuda, tmp_path, kind):
    from particleclem_amd.hydro import DeviceLevel, HydroContext
    w = (wl.premixed(n=32, box=16, nsteps=0) if kind == "premixed"
         else wl.sod(n=16, max_grid=8, nsteps=0))
    ctx = HydroContext(w.nspec)
    lev = DeviceLevel.from_workload(ctx, w)
    lev.init_dt()
    lev.step(2)
    path = tmp_path / f"plt_{kind}"
    lev.write_plotfile(path, "ALL", time=1.5e-3, step=2, prob_lo=(0.5, 0.0, -1.0))
    p = pf.read_plotfile(str(path))
    assert p.names == all_names(w.nspec)
    assert p.step == 2 and p.time == 1.5e-3 and p.prob_lo == (0.5, 0.0, -1.0)
    assert tuple(p.dx) == tuple(w.dx)
    fabs = [lev.read_box(b) for b in range(len(w.boxes))]
    H = ob.HostLevel(w.nspec, w.boxes, w.domlo, w.domhi, w.bclo, w.bchi, w.dx)
    for b, f in enumerate(fabs):
        H.fab(b)[...] = f
    H.fill_ghosts()
    codes = [code_of(n, w.nspec) for n in p.names]
    worst = 0.0
    nbits = 0
    for b, (lo, hi) in enumerate(w.boxes):
        assert p.boxes[b] == (tuple(lo), tuple(hi))
        got = p.fabs[b]
        nst = 7 + w.nspec
        np.testing.assert_array_equal(got[:nst], fabs[b][VALID])
        U = np.ascontiguousarray(H.fab(b)[:, 3:-3, 3:-3, 3:-3])
        ref = derive(w.nspec, U, lo, hi, [c for c, _ in codes], [m for _, m in codes], w.dx)
        e, nb, _ = compare(got, ref)
        worst = max(worst, e)
        nbits += nb
    print(f"{kind}: {len(p.names)} plot variables, max rel err vs oracle {worst:.2e}, "
          f"{nbits} values not bitwise")
    assert worst <= 1e-12
    # defaults and errors
    lev.write_plotfile(tmp_path / "plt_state")
    assert pf.read_plotfile(str(tmp_path / "plt_state")).names == all_names(w.nspec)[:7 + w.nspec]
    from particleclem_amd._lib import HydroError
    with pytest.raises(HydroError):
        lev.write_plotfile(tmp_path / "plt_bad", "density nosuchvar")
    lev.close()
    ctx.close()
