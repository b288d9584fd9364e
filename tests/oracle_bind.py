"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: the oracle is the checker, never the thing measured.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from particleclem_amd.ctypes_types import (Array4, Box, LevelDesc, Params,
                                           a4_from_numpy, make_box)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# PC_ORACLE_SO names another build of the oracle (the planted-mistake check
# in tests/test_oracle_mutants.py loads deliberately broken copies this way)
ORACLE_SO = os.environ.get("PC_ORACLE_SO") or os.path.join(ROOT, "oracle", "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle")])
    L = C.CDLL(ORACLE_SO)
    P = C.POINTER
    d3 = P(C.c_double)
    i3 = P(C.c_int)
    L.pc_ref_params_default.argtypes = [P(Params)]
    L.pc_ref_ctoprim.argtypes = [C.c_int, Box, Array4, Array4, Array4]
    L.pc_ref_srctoprim.argtypes = [C.c_int, Box, Array4, Array4, Array4, Array4]
    L.pc_ref_umdrv.argtypes = [C.c_int, C.c_double, Box, i3, i3, i3, i3,
                               Array4, Array4, Array4, Array4, Array4, d3,
                               C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_int, C.c_double, P(Array4), P(Array4),
                               Array4, d3]
    L.pc_ref_hydro_source_box.argtypes = [C.c_int, Box, i3, i3, i3, i3, Array4,
                                          Array4, Array4, d3, C.c_double,
                                          P(Params), P(Array4), P(Array4), Array4]
    L.pc_ref_level_arena_size.argtypes = [C.c_int, P(LevelDesc), P(C.c_longlong)]
    L.pc_ref_level_arena_size.restype = C.c_longlong
    L.pc_ref_level_fill_ghosts.argtypes = [C.c_int, P(LevelDesc), d3]
    L.pc_ref_level_estdt.argtypes = [C.c_int, P(LevelDesc), P(Params), d3, d3]
    L.pc_ref_level_step.argtypes = [C.c_int, P(LevelDesc), P(Params), d3,
                                    C.c_int, d3, d3, C.c_int]
    L.pc_ref_level_init_dt.argtypes = [C.c_int, P(LevelDesc), P(Params), d3,
                                       C.c_double, d3]
    L.pc_ref_transdo_point.argtypes = [C.c_int, C.c_int, C.c_int] + [d3] * 6 + [
        C.c_double, C.c_double, d3, d3]
    L.pc_ref_transdd_point.argtypes = [C.c_int, C.c_int] + [d3] * 10 + [
        C.c_double, d3, C.c_double, C.c_double, C.c_double, d3, d3]
    L.pc_ref_pdivu_point.argtypes = [d3, d3, d3, C.c_double]
    L.pc_ref_pdivu_point.restype = C.c_double
    L.pc_ref_prim_to_cons.argtypes = [C.c_int, Box, Array4, C.c_int, Array4]
    L.pc_ref_eos.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double,
                             C.c_double, d3, d3]
    L.pc_ref_riemann.argtypes = [C.c_int, d3, d3, d3, d3, C.c_double, d3]
    L.pc_ref_ppm_reconstruct.argtypes = [d3, C.c_double, d3]
    L.pc_ref_ppm_int_profile.argtypes = [C.c_double] * 6 + [d3, d3]
    L.pc_ref_plm_slope_line.argtypes = [d3, C.c_double, C.c_int]
    L.pc_ref_plm_slope_line.restype = C.c_double
    L.pc_ref_mol_src_box.argtypes = [C.c_int, Box, Array4, Array4, P(Params),
                                     P(Array4), P(Array4), Array4]
    L.pc_ref_laxf.argtypes = [C.c_int, d3, d3, d3, d3, d3]
    L.pc_ref_derive.argtypes = [C.c_int, Box, Array4, i3, i3, C.c_int, Array4, d3]
    L.pc_ref_fluxreg_create.argtypes = [C.c_int, P(Box), C.c_int, P(Box), C.c_int, i3,
                                        i3, i3, C.c_int]
    L.pc_ref_fluxreg_create.restype = C.c_void_p
    L.pc_ref_fluxreg_destroy.argtypes = [C.c_void_p]
    L.pc_ref_fluxreg_nfaces.argtypes = [C.c_void_p]
    L.pc_ref_fluxreg_crse_add.argtypes = [C.c_void_p, Box, P(Array4), C.c_double]
    L.pc_ref_fluxreg_fine_add.argtypes = [C.c_void_p, Box, P(Array4), C.c_double]
    L.pc_ref_fluxreg_reflux.argtypes = [C.c_void_p, Box, Array4, C.c_double]
    L.pc_ref_average_down.argtypes = [C.c_int, Box, Array4, Array4, C.c_int]
    _lib = L
    return L


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iarr(v):
    return (C.c_int * 3)(*v)


def darr(v):
    return (C.c_double * 3)(*v)


def eos(kind, which, a=0.0, b=0.0, c=0.0, Y=None, nout=1):
    out = np.zeros(max(nout, 9))
    Yp = None
    if Y is not None:
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        Yp = dptr(Y)
    rc = lib().pc_ref_eos(kind, which, a, b, c, Yp, dptr(out))
    assert rc == 0
    return out[:nout] if nout > 1 else out[0]


def riemann(kind, left5, spl, right5, spr, cav=1.0):
    out = np.zeros(12)
    l5 = np.ascontiguousarray(left5, dtype=np.float64)
    r5 = np.ascontiguousarray(right5, dtype=np.float64)
    sl = np.ascontiguousarray(spl, dtype=np.float64)
    sr = np.ascontiguousarray(spr, dtype=np.float64)
    rc = lib().pc_ref_riemann(kind, dptr(l5), dptr(sl), dptr(r5), dptr(sr),
                              cav, dptr(out))
    assert rc == 0
    keys = ["ustar", "f_rho", "f_un", "f_ut", "f_utt", "f_eden", "f_eint",
            "u_gd", "ut_gd", "utt_gd", "p_gd", "game_gd"]
    return dict(zip(keys, out))


def laxf(kind, left5, spl, right5, spr):
    out = np.zeros(8 + kind)
    l5 = np.ascontiguousarray(left5, dtype=np.float64)
    r5 = np.ascontiguousarray(right5, dtype=np.float64)
    sl = np.ascontiguousarray(spl, dtype=np.float64)
    sr = np.ascontiguousarray(spr, dtype=np.float64)
    assert lib().pc_ref_laxf(kind, dptr(l5), dptr(sl), dptr(r5), dptr(sr),
                             dptr(out)) == 0
    keys = ["ustar", "maxeig", "f_rho", "f_un", "f_ut", "f_utt", "f_eden",
            "f_eint"]
    d = dict(zip(keys, out[:8]))
    d["f_rhoY"] = out[8:].copy()
    return d


def mol_src_box(kind, S, lo, hi, dx, prm, want_flux=False):
    """MOL source term of one box: S is the (NVAR, nz+8, ny+8, nx+8) state
    with 4 ghosts around [lo, hi]; returns molsrc on the box (and the three
    extensive face-flux arrays when want_flux)."""
    nvar = 7 + kind
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    S = np.ascontiguousarray(S, dtype=np.float64)
    glo = [x - 4 for x in lo]
    src = np.zeros((nvar, n[2], n[1], n[0]))
    vol = np.full((1, n[2], n[1], n[0]), dx[0] * dx[1] * dx[2])
    area_v = [dx[1] * dx[2], dx[2] * dx[0], dx[0] * dx[1]]
    areas, flx = [], []
    for d in range(3):
        m = list(n)
        m[d] += 1
        areas.append(np.full((1, m[2], m[1], m[0]), area_v[d]))
        flx.append(np.zeros((nvar, m[2], m[1], m[0])))
    A = (Array4 * 3)(*[a4_from_numpy(a, lo) for a in areas])
    F = (Array4 * 3)(*[a4_from_numpy(f, lo) for f in flx])
    rc = lib().pc_ref_mol_src_box(kind, make_box(lo, hi), a4_from_numpy(S, glo),
                                  a4_from_numpy(src, lo), C.byref(prm),
                                  F if want_flux else None, A,
                                  a4_from_numpy(vol, lo))
    assert rc == 0, rc
    return (src, flx) if want_flux else src


def params(**kw):
    from particleclem_amd.ctypes_types import make_params
    return make_params(**kw)


class HostLevel:
    """A level (box table + host arena) run by the oracle."""

    def __init__(self, kind, boxes, domlo, domhi, bclo, bchi, dx,
                 periodic_self=False, problo=(0.0, 0.0, 0.0)):
        self.kind = kind
        self.nvar = 7 + kind
        self.boxes = [(list(lo), list(hi)) for lo, hi in boxes]
        self._boxarr = (Box * len(boxes))(*[make_box(lo, hi) for lo, hi in boxes])
        self.desc = LevelDesc()
        self.desc.nboxes = len(boxes)
        self.desc.boxes = C.cast(self._boxarr, C.POINTER(Box))
        self.desc.domlo = iarr(domlo)
        self.desc.domhi = iarr(domhi)
        self.desc.bclo = iarr(bclo)
        self.desc.bchi = iarr(bchi)
        self.desc.periodic_self = 1 if periodic_self else 0
        self.desc.dx = darr(dx)
        self.desc.problo = darr(problo)
        offs = (C.c_longlong * len(boxes))()
        n = lib().pc_ref_level_arena_size(kind, C.byref(self.desc), offs)
        self.offsets = list(offs)
        self.arena = np.zeros(n, dtype=np.float64)

    def fab(self, b):
        """numpy view (ncomp, nz, ny, nx) of box b's FAB (4 ghosts)."""
        lo, hi = self.boxes[b]
        s = [hi[d] - lo[d] + 1 + 8 for d in range(3)]
        n = s[0] * s[1] * s[2] * self.nvar
        o = self.offsets[b]
        return self.arena[o:o + n].reshape(self.nvar, s[2], s[1], s[0])

    def valid(self, b):
        return self.fab(b)[:, 4:-4, 4:-4, 4:-4]

    def set_prims(self, workload):
        """Conservative valid-cell state from the workload's primitives."""
        for b, (lo, hi) in enumerate(self.boxes):
            P = workload.prims[b]
            fab = self.fab(b)
            U = np.zeros((self.nvar,) + P.shape[1:])
            rc = lib().pc_ref_prim_to_cons(self.kind, make_box(lo, hi),
                                           a4_from_numpy(P, lo), workload.prim_mode,
                                           a4_from_numpy(U, lo))
            assert rc == 0
            fab[:, 4:-4, 4:-4, 4:-4] = U

    def fill_ghosts(self):
        assert lib().pc_ref_level_fill_ghosts(self.kind, C.byref(self.desc),
                                              dptr(self.arena)) == 0

    def estdt(self, prm):
        out = C.c_double(0.0)
        assert lib().pc_ref_level_estdt(self.kind, C.byref(self.desc),
                                        C.byref(prm), dptr(self.arena),
                                        C.byref(out)) == 0
        return out.value

    time = 0.0

    def init_dt(self, prm):
        """computeInitialDt at the level's time (init_shrink, stop_time)."""
        out = C.c_double(0.0)
        assert lib().pc_ref_level_init_dt(self.kind, C.byref(self.desc),
                                          C.byref(prm), dptr(self.arena),
                                          self.time, C.byref(out)) == 0
        return out.value

    def step(self, prm, nsteps, dt, nthreads=0):
        d = C.c_double(dt)
        t = C.c_double(self.time)
        rc = lib().pc_ref_level_step(self.kind, C.byref(self.desc), C.byref(prm),
                                     dptr(self.arena), nsteps, C.byref(d),
                                     C.byref(t), nthreads)
        assert rc == 0, rc
        self.time = t.value
        return d.value


def transdo(kind, dir, t, qplus, qminus, fl_lo, fl_hi, qi_lo, qi_hi, gamc,
            cdtdx):
    """pc_transdo on one cell: returns the corrected (plus, minus) states."""
    qv = 8 + kind
    op, om = np.zeros(qv), np.zeros(qv)
    a = [np.ascontiguousarray(x, dtype=np.float64)
         for x in (qplus, qminus, fl_lo, fl_hi, qi_lo, qi_hi)]
    assert lib().pc_ref_transdo_point(kind, dir, t, *[dptr(x) for x in a], gamc,
                                      cdtdx, dptr(op), dptr(om)) == 0
    return op, om


def transdd(kind, dir, qplus, qminus, f0, f1, qi0, qi1, gamc, srcq, hdt, c0,
            c1):
    """pc_transdd on one cell; f0/f1/qi0/qi1 are (lo, hi) pairs."""
    qv = 8 + kind
    op, om = np.zeros(qv), np.zeros(qv)
    a = [np.ascontiguousarray(x, dtype=np.float64)
         for x in (qplus, qminus, f0[0], f0[1], f1[0], f1[1], qi0[0], qi0[1],
                   qi1[0], qi1[1])]
    s = np.ascontiguousarray(srcq if srcq is not None else np.zeros(qv),
                             dtype=np.float64)
    assert lib().pc_ref_transdd_point(kind, dir, *[dptr(x) for x in a], gamc,
                                      dptr(s), hdt, c0, c1, dptr(op),
                                      dptr(om)) == 0
    return op, om


def pdivu(pg, un, area, vol):
    """pc_pdivu of one cell from its six faces' (p_gd, u_gd, area)."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (pg, un, area)]
    return lib().pc_ref_pdivu_point(*[dptr(x) for x in a], vol)
