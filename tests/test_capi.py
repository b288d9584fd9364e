"""The C ABI without a GPU: the library loads, exports every entry point
include/pc_hydro.h declares, fills the reference's parameter defaults and
rejects bad arguments with status codes (no kernel is launched)."""
import ctypes as C
import os
import re

import pytest

from particleclem_amd import _lib
from particleclem_amd.ctypes_types import (Array4, Params, make_box,
                                           make_params, null_a4)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "pc_hydro.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pc_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == names


def test_params_default_match_reference():
    # Source/Params/_cpp_parameters:53,96,103,113,118-130,267-289
    p = Params()
    _lib.lib().pc_params_default(C.byref(p))
    d = make_params()
    for f, _ in Params._fields_:
        assert getattr(p, f) == getattr(d, f), f
    assert p.ppm_type == 0 and p.plm_iorder == 4 and p.difmag == 0.1
    assert p.cfl == 0.8 and p.change_max == 1.1 and p.max_dt == 1e200


def _ctx(kind=1):
    h = C.c_void_p()
    assert _lib.lib().pc_ctx_create(kind, C.byref(h)) == 0
    return h


def test_context_queries():
    L = _lib.lib()
    for kind, nv, qv in [(1, 8, 9), (9, 16, 17)]:
        h = _ctx(kind)
        assert L.pc_ctx_num_species(h) == kind
        assert L.pc_ctx_nvar(h) == nv
        assert L.pc_ctx_qvar(h) == qv
        assert L.pc_ctx_destroy(h) == 0
    h = C.c_void_p()
    assert L.pc_ctx_create(3, C.byref(h)) == -1  # PC_ERR_ARG


def _a4(lo, hi, nc):
    a = Array4()
    a.p = 0x1000  # never dereferenced: validation fails before any launch
    nx, ny, nz = (hi[d] - lo[d] + 1 for d in range(3))
    a.jstride, a.kstride, a.nstride = nx, nx * ny, nx * ny * nz
    a.begin = (C.c_int * 3)(*lo)
    a.end = (C.c_int * 3)(hi[0] + 1, hi[1] + 1, hi[2] + 1)
    a.ncomp = nc
    return a


def _umdrv(h, ppm_type=1, plm_iorder=4, weno=0, grow=4, nc=8):
    L = _lib.lib()
    lo, hi = [0, 0, 0], [7, 7, 7]
    g = [x - grow for x in lo], [x + grow for x in hi]
    i3 = lambda v: (C.c_int * 3)(*v)
    flx = (Array4 * 3)(*[_a4(lo, [hi[0] + (d == 0), hi[1] + (d == 1), hi[2] + (d == 2)], nc)
                         for d in range(3)])
    ar = (Array4 * 3)(*[_a4(lo, [hi[0] + (d == 0), hi[1] + (d == 1), hi[2] + (d == 2)], 1)
                        for d in range(3)])
    return L.pc_umdrv_v1(h, 0.0, make_box(lo, hi), i3([0, 0, 0]), i3([7, 7, 7]),
                         i3([0, 0, 0]), i3([0, 0, 0]), _a4(*g, nc), _a4(lo, hi, nc),
                         _a4(*g, nc + 1), _a4(*g, 6), null_a4(),
                         (C.c_double * 3)(0.1, 0.1, 0.1), 1e-3, ppm_type,
                         plm_iorder, 1, weno, 1, 0.1, flx, ar, _a4(lo, hi, 1),
                         None, None)


def test_umdrv_rejects_bad_options_with_codes():
    h = _ctx(1)
    assert _umdrv(h, ppm_type=2) == -3          # PC_ERR_PPM_TYPE
    assert _umdrv(h, plm_iorder=3) == -4        # PC_ERR_PLM_IORDER
    assert _umdrv(h, weno=1) == -7              # PC_ERR_UNSUPPORTED
    assert _umdrv(h, grow=3) == -2              # PC_ERR_GROW
    assert _umdrv(h, nc=5) == -1                # PC_ERR_ARG (ncomp)
    msg = _lib.lib().pc_last_error(h).decode()
    assert "ncomp" in msg
    _lib.lib().pc_ctx_destroy(h)


def test_level_rejects_non_tiling_boxes():
    from particleclem_amd.ctypes_types import Box, LevelDesc
    L = _lib.lib()
    h = _ctx(1)
    boxes = (Box * 1)(make_box([0, 0, 0], [7, 7, 7]))
    d = LevelDesc()
    d.nboxes = 1
    d.boxes = C.cast(boxes, C.POINTER(Box))
    d.domlo = (C.c_int * 3)(0, 0, 0)
    d.domhi = (C.c_int * 3)(3, 3, 3)   # box outside the domain
    d.dx = (C.c_double * 3)(0.1, 0.1, 0.1)
    lev = C.c_void_p()
    p = make_params(ppm_type=5)
    assert L.pc_level_create_v1(h, C.byref(d), C.byref(p), C.byref(lev)) == -3
    p = make_params(ppm_type=1)
    assert L.pc_level_create_v1(h, C.byref(d), C.byref(p), C.byref(lev)) == -1
    L.pc_ctx_destroy(h)


def test_empty_inputs_are_rejected_with_codes():
    """Empty boxes and empty box tables are argument errors (PC_ERR_ARG),
    caught before any device work; a batch step over no requests and the
    branch-stream switches reject null handles and bad values."""
    from particleclem_amd.ctypes_types import Box, LevelDesc
    L = _lib.lib()
    h = _ctx(1)
    i3 = lambda v: (C.c_int * 3)(*v)
    lo, hi = [0, 0, 0], [-1, 7, 7]          # empty in x
    flx = (Array4 * 3)(null_a4(), null_a4(), null_a4())
    rc = L.pc_hydro_source_box_v1(h, make_box(lo, hi), i3([0, 0, 0]), i3([7, 7, 7]),
                                  i3([0, 0, 0]), i3([0, 0, 0]), null_a4(), null_a4(),
                                  null_a4(), (C.c_double * 3)(0.1, 0.1, 0.1), 1e-3,
                                  C.byref(make_params(ppm_type=1)), flx, flx, null_a4(),
                                  None)
    assert rc == -1
    assert "empty box" in L.pc_last_error(h).decode()
    d = LevelDesc()
    d.nboxes = 0
    d.domlo = i3([0, 0, 0])
    d.domhi = i3([7, 7, 7])
    d.dx = (C.c_double * 3)(0.1, 0.1, 0.1)
    lev = C.c_void_p()
    p = make_params(ppm_type=1)
    assert L.pc_level_create_v1(h, C.byref(d), C.byref(p), C.byref(lev)) == -1
    boxes = (Box * 2)(make_box([0, 0, 0], [7, 7, 7]), make_box([8, 0, 0], [7, 7, 7]))
    d.nboxes = 2
    d.boxes = C.cast(boxes, C.POINTER(Box))
    assert L.pc_level_create_v1(h, C.byref(d), C.byref(p), C.byref(lev)) == -1
    assert "empty box 1" in L.pc_last_error(h).decode()
    assert L.pc_batch_step_v1(None, None, 0, None) == -1
    assert L.pc_level_set_branches_v1(None, 1) == -1
    assert L.pc_batch_set_branches_v1(None, 1) == -1
    L.pc_ctx_destroy(h)


def test_boxes_beyond_32bit_offsets_are_rejected():
    """Kernels address a FAB with 32-bit offsets (buffer loads take the
    byte offset of component and cell as one 32-bit value): a level box
    whose stage arrays would exceed 2 GiB (248^3 with 9 species: q alone is
    256^3 x 17 doubles = 2.28 GB) and a pc_umdrv_v1 call with a q FAB over
    2 GiB fail with PC_ERR_UNSUPPORTED before any device work."""
    from particleclem_amd.ctypes_types import Box, LevelDesc
    L = _lib.lib()
    h = _ctx(9)
    i3 = lambda v: (C.c_int * 3)(*v)
    d = LevelDesc()
    d.nboxes = 1
    for n, want in ((248, -7), (200, None)):
        boxes = (Box * 1)(make_box([0, 0, 0], [n - 1] * 3))
        d.boxes = C.cast(boxes, C.POINTER(Box))
        d.domlo = i3([0, 0, 0])
        d.domhi = i3([n - 1] * 3)
        d.dx = (C.c_double * 3)(0.1, 0.1, 0.1)
        lev = C.c_void_p()
        p = make_params(ppm_type=1)
        if want is None:
            continue  # 200^3 passes the size check (allocation needs a GPU)
        assert L.pc_level_create_v1(h, C.byref(d), C.byref(p), C.byref(lev)) == want
        assert "32-bit offsets" in L.pc_last_error(h).decode()
    # drop-in: a 16^3 box whose caller q FAB spans 17 x 2^27 doubles
    lo, hi = [0, 0, 0], [15, 15, 15]
    g = [-4] * 3, [19] * 3
    big_q = _a4(*g, 17)
    big_q.nstride = 1 << 27
    flx = (Array4 * 3)(*[_a4(lo, [hi[0] + (k == 0), hi[1] + (k == 1), hi[2] + (k == 2)], 16)
                         for k in range(3)])
    ar = (Array4 * 3)(*[_a4(lo, [hi[0] + (k == 0), hi[1] + (k == 1), hi[2] + (k == 2)], 1)
                        for k in range(3)])
    rc = L.pc_umdrv_v1(h, 0.0, make_box(lo, hi), i3([0, 0, 0]), i3([15] * 3),
                       i3([0, 0, 0]), i3([0, 0, 0]), _a4(*g, 16), _a4(lo, hi, 16),
                       big_q, _a4(*g, 6), null_a4(), (C.c_double * 3)(0.1, 0.1, 0.1),
                       1e-3, 1, 4, 1, 0, 1, 0.1, flx, ar, _a4(lo, hi, 1), None, None)
    assert rc == -7, rc
    assert "32-bit offsets" in L.pc_last_error(h).decode()
    L.pc_ctx_destroy(h)


def test_params_carry_stop_time():
    p = Params()
    _lib.lib().pc_params_default(C.byref(p))
    assert p.stop_time == -1.0
