"""The device flux register and average-down (pc_fluxreg_*_v1,
pc_average_down_v1) against the oracle's dense restatement on the same
inputs: the two-level run's coarse and fine Godunov fluxes are fed to both,
the refluxed coarse boxes and the averaged-down block must be bitwise
equal, and the register face count must agree."""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob
from particleclem_amd.ctypes_types import Array4, a4_from_numpy, make_box
from test_oracle_fluxreg import CB, FB, FREG, NC, RATIO, boxarr, ref_reg, two_level_run

pytestmark = pytest.mark.gpu


def test_device_register_equals_oracle(cuda):
    import torch
    from particleclem_amd.hydro import HydroContext, a4_from_tensor
    U0, want, calls, Uc1, Uf = two_level_run(reflux=True)
    ctx = HydroContext(1)
    L = ctx.L
    reg = C.c_void_p()
    ctx._chk(L.pc_fluxreg_create_v1(ctx.h, boxarr(CB), len(CB), boxarr(FB), len(FB),
                                    ob.iarr([0] * 3), ob.iarr([NC - 1] * 3),
                                    ob.iarr([1, 1, 1]), RATIO, C.byref(reg)))
    r = ref_reg(8, CB, FB, [0] * 3, [NC - 1] * 3, [1, 1, 1])
    assert L.pc_fluxreg_nfaces_v1(reg) == ob.lib().pc_ref_fluxreg_nfaces(r)
    ob.lib().pc_ref_fluxreg_destroy(r)
    ctx._chk(L.pc_fluxreg_reset_v1(reg, None))
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    keep = []
    for kind, b, lo, hi, flx, scale in calls:
        d = [T(f) for f in flx]
        keep.append(d)
        fl = (Array4 * 3)(*[a4_from_tensor(x, lo) for x in d])
        fn = L.pc_fluxreg_crse_add_v1 if kind == "crse" else L.pc_fluxreg_fine_add_v1
        ctx._chk(fn(reg, b, make_box(lo, hi), fl, scale, None))
    out = Uc1.copy()
    vol = (1.0 / NC) ** 3
    for b, (lo, hi) in enumerate(CB):
        s = (slice(None), slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1), slice(lo[0], hi[0] + 1))
        box = T(out[s])
        ctx._chk(L.pc_fluxreg_reflux_v1(reg, b, make_box(lo, hi), a4_from_tensor(box, lo),
                                        Array4(), vol, None))
        torch.cuda.synchronize()
        out[s] = box.cpu().numpy()
    flo = RATIO * FREG[0][0]
    crse = T(out[:, 4:12, 4:12, 4:12])
    ctx._chk(L.pc_average_down_v1(ctx.h, make_box(*FREG), a4_from_tensor(crse, FREG[0]),
                                  a4_from_tensor(T(Uf), [flo] * 3), RATIO, None))
    torch.cuda.synchronize()
    out[:, 4:12, 4:12, 4:12] = crse.cpu().numpy()
    nd = int(np.count_nonzero(out != want))
    print(f"two-level reflux + average-down: {nd} of {out.size} values differ from the oracle")
    assert nd == 0
    # the register changed the cells next to the refined block only
    changed = np.any(out != Uc1, axis=0)
    changed[4:12, 4:12, 4:12] = False
    ring = np.zeros_like(changed)
    ring[3:13, 3:13, 3:13] = True
    ring[4:12, 4:12, 4:12] = False
    assert changed.any() and not (changed & ~ring).any()
    ctx._chk(L.pc_fluxreg_destroy_v1(reg))
    ctx.close()
