"""Host-side mirror of the reference's hydro interface over the C ABI.

  HydroContext.ctoprim / srctoprim / umdrv / hydro_source_box
      -> pc_ctoprim, pc_srctoprim, pc_umdrv (Source/Hydro.H:910-935) and one
         box of PeleC::construct_hydro_source (Source/Hydro.cpp:73-252)
  HydroContext.mol_src_box
      -> one box of PeleC::getMOLSrcTerm, hydro only (Source/Diffusion.cpp:90-654)
  DeviceLevel.step / init_dt / estdt / fill_ghosts
      -> do_sdc_iteration's hydro path + computeTemp + computeNewDt
         (Source/Advance.cpp:266-416, Source/PeleC.cpp:811-1034), or with
         do_mol=1 do_mol_advance (Source/Advance.cpp:49-227)

Device arrays are torch tensors (HIP memory) shaped (ncomp, nz, ny, nx),
i.e. AMReX FAB layout; `lo` gives the index of element [.., 0, 0, 0].
Errors from the library raise HydroError; there is no CPU path here.
"""
import ctypes as C

import numpy as np

from . import _lib
from .ctypes_types import (Array4, Box, LevelDesc, a4_from_ptr, make_box,
                           make_params, null_a4)

check = _lib.check


def _stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def a4_from_tensor(t, lo):
    """pc_array4_t of a contiguous float64 tensor (ncomp, nz, ny, nx)."""
    assert t.dtype.is_floating_point and t.element_size() == 8
    assert t.is_contiguous()
    nc, nz, ny, nx = t.shape
    hi = [lo[0] + nx - 1, lo[1] + ny - 1, lo[2] + nz - 1]
    return a4_from_ptr(t.data_ptr(), list(lo), hi, nc)


def _i3(v):
    return (C.c_int * 3)(*v)


def _d3(v):
    return (C.c_double * 3)(*v)


class HydroContext:
    """An EOS bound to the current HIP device (pc_ctx_t)."""

    def __init__(self, eos_kind):
        self.L = _lib.lib()
        h = C.c_void_p()
        check(self.L.pc_ctx_create(int(eos_kind), C.byref(h)))
        self.h = h
        self.nspec = self.L.pc_ctx_num_species(h)
        self.nvar = self.L.pc_ctx_nvar(h)
        self.qvar = self.L.pc_ctx_qvar(h)

    def close(self):
        if self.h:
            self.L.pc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        return check(rc, self.h)

    def ctoprim(self, qbx, u, q, qaux, stream=None):
        """qbx = (lo, hi); u, q, qaux: (tensor, lo) pairs."""
        self._chk(self.L.pc_ctoprim_v1(self.h, make_box(*qbx), a4_from_tensor(*u),
                                       a4_from_tensor(*q), a4_from_tensor(*qaux),
                                       _stream_ptr(stream)))

    def srctoprim(self, qbx, q, qaux, src, srcq, stream=None):
        self._chk(self.L.pc_srctoprim_v1(self.h, make_box(*qbx), a4_from_tensor(*q),
                                         a4_from_tensor(*qaux), a4_from_tensor(*src),
                                         a4_from_tensor(*srcq), _stream_ptr(stream)))

    def umdrv(self, bx, domlo, domhi, bclo, bchi, uin, uout, q, qaux, src_q,
              dx, dt, ppm_type, plm_iorder, use_flattening, difmag, flx, a,
              vol, use_hybrid_weno=0, weno_scheme=1, stream=None):
        """pc_umdrv with the reference's argument order; array arguments are
        (tensor, lo) pairs, src_q may be None (zero sources)."""
        fl = (Array4 * 3)(*[a4_from_tensor(*f) for f in flx])
        aa = (Array4 * 3)(*[a4_from_tensor(*x) for x in a])
        sq = a4_from_tensor(*src_q) if src_q is not None else null_a4()
        self._chk(self.L.pc_umdrv_v1(
            self.h, 0.0, make_box(*bx), _i3(domlo), _i3(domhi), _i3(bclo),
            _i3(bchi), a4_from_tensor(*uin), a4_from_tensor(*uout),
            a4_from_tensor(*q), a4_from_tensor(*qaux), sq, _d3(dx), float(dt),
            int(ppm_type), int(plm_iorder), int(use_flattening),
            int(use_hybrid_weno), int(weno_scheme), float(difmag), fl, aa,
            a4_from_tensor(*vol), None, _stream_ptr(stream)))

    def hydro_source_box(self, bx, domlo, domhi, bclo, bchi, sarr, src,
                         hyd_src, dx, dt, params, flx, a, vol, stream=None):
        fl = (Array4 * 3)(*[a4_from_tensor(*f) for f in flx])
        aa = (Array4 * 3)(*[a4_from_tensor(*x) for x in a])
        s = a4_from_tensor(*src) if src is not None else null_a4()
        prm = params if not isinstance(params, dict) else make_params(**params)
        self._chk(self.L.pc_hydro_source_box_v1(
            self.h, make_box(*bx), _i3(domlo), _i3(domhi), _i3(bclo), _i3(bchi),
            a4_from_tensor(*sarr), s, a4_from_tensor(*hyd_src), _d3(dx),
            float(dt), C.byref(prm), fl, aa, a4_from_tensor(*vol),
            _stream_ptr(stream)))

    def mol_src_box(self, bx, sarr, molsrc, params, a, vol, flx=None,
                    stream=None):
        """getMOLSrcTerm of one box (pc_mol_src_box_v1): molsrc = -div(F)/vol
        of the MOL fluxes of sarr; flx (three (tensor, lo) pairs) or None."""
        aa = (Array4 * 3)(*[a4_from_tensor(*x) for x in a])
        fl = (Array4 * 3)(*[a4_from_tensor(*f) for f in flx]) if flx else None
        prm = params if not isinstance(params, dict) else make_params(**params)
        self._chk(self.L.pc_mol_src_box_v1(
            self.h, make_box(*bx), a4_from_tensor(*sarr),
            a4_from_tensor(*molsrc), C.byref(prm), fl, aa,
            a4_from_tensor(*vol), _stream_ptr(stream)))


class DeviceLevel:
    """One AMR level (a box table) resident on the GPU (pc_level_t)."""

    def __init__(self, ctx, boxes, domlo, domhi, bclo, bchi, dx,
                 periodic_self=False, params=None, problo=(0.0, 0.0, 0.0)):
        self.ctx = ctx
        self.L = ctx.L
        self.boxes = [(list(lo), list(hi)) for lo, hi in boxes]
        self._boxarr = (Box * len(boxes))(*[make_box(lo, hi) for lo, hi in boxes])
        d = LevelDesc()
        d.nboxes = len(boxes)
        d.boxes = C.cast(self._boxarr, C.POINTER(Box))
        d.domlo = _i3(domlo)
        d.domhi = _i3(domhi)
        d.bclo = _i3(bclo)
        d.bchi = _i3(bchi)
        d.periodic_self = 1 if periodic_self else 0
        d.dx = _d3(dx)
        d.problo = _d3(problo)
        self.desc = d
        self.params = make_params(**(params or {}))
        h = C.c_void_p()
        check(self.L.pc_level_create_v1(ctx.h, C.byref(d), C.byref(self.params),
                                        C.byref(h)), ctx.h)
        self.h = h

    @classmethod
    def from_workload(cls, ctx, w, **overrides):
        prm = dict(w.params)
        prm.update(overrides)
        lev = cls(ctx, w.boxes, w.domlo, w.domhi, w.bclo, w.bchi, w.dx,
                  periodic_self=w.periodic_self, params=prm)
        lev.load_prims(w)
        return lev

    def close(self):
        if self.h:
            self.L.pc_level_destroy_v1(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        return check(rc, self.ctx.h)

    def fab_shape(self, b):
        lo, hi = self.boxes[b]
        return (self.ctx.nvar,) + tuple(hi[d] - lo[d] + 9 for d in (2, 1, 0))

    def load_prims(self, w, stream=None):
        """Valid cells of every box from the workload's primitive fields."""
        import torch
        for b, (lo, hi) in enumerate(self.boxes):
            P = torch.from_numpy(np.ascontiguousarray(w.prims[b])).to("cuda")
            self._chk(self.L.pc_level_prim_to_cons_v1(
                self.h, b, a4_from_tensor(P, lo), int(w.prim_mode),
                _stream_ptr(stream)))
        torch.cuda.synchronize()

    def read_box(self, b):
        out = np.empty(self.fab_shape(b))
        self._chk(self.L.pc_level_read_box_v1(self.h, b, out.ctypes.data))
        return out

    def write_box(self, b, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        assert arr.shape == self.fab_shape(b)
        self._chk(self.L.pc_level_write_box_v1(self.h, b, arr.ctypes.data))

    def set_dt(self, dt):
        self._chk(self.L.pc_level_set_dt_v1(self.h, float(dt)))

    def get_dt(self):
        v = C.c_double()
        self._chk(self.L.pc_level_get_dt_v1(self.h, C.byref(v)))
        return v.value

    def set_time(self, t):
        self._chk(self.L.pc_level_set_time_v1(self.h, float(t)))

    def get_time(self):
        v = C.c_double()
        self._chk(self.L.pc_level_get_time_v1(self.h, C.byref(v)))
        return v.value

    def init_dt(self, stream=None):
        self._chk(self.L.pc_level_init_dt_v1(self.h, _stream_ptr(stream)))

    def estdt(self, stream=None):
        v = C.c_double()
        self._chk(self.L.pc_level_estdt_v1(self.h, C.byref(v), _stream_ptr(stream)))
        return v.value

    def fill_ghosts(self, stream=None):
        self._chk(self.L.pc_level_fill_ghosts_v1(self.h, _stream_ptr(stream)))

    def step(self, nsteps=1, stream=None):
        self._chk(self.L.pc_level_step_v1(self.h, int(nsteps), _stream_ptr(stream)))

    def status(self):
        bad = C.c_int()
        small = C.c_int()
        rc = self.L.pc_level_status_v1(self.h, C.byref(bad), C.byref(small))
        return rc, bad.value, small.value

    def autotune(self, reps=3, stream=None):
        """Pick the fastest launch shape of every stage on this level
        (synchronous; state and dt unchanged)."""
        self._chk(self.L.pc_level_autotune_v1(self.h, int(reps), _stream_ptr(stream)))

    def plans(self):
        """{stage: (shape (lx, ly, lz), 'runs' or 'tiles' (the LDS tiles of
        the PPM trace and of the FP64 MOL fluxes), workgroups, logical blocks,
        tuned ms)}"""
        out = {}
        for sid, name in enumerate(_lib.STAGES):
            sh = (C.c_int * 5)()
            ms = C.c_double()
            self._chk(self.L.pc_level_plan_v1(self.h, sid, sh, C.byref(ms)))
            shape = ("runs" if sh[0] == -1 else "tiles" if sh[0] < 0
                     else (1 << sh[0], 1 << sh[1], 1 << sh[2]))
            out[name] = (shape, sh[3], sh[4], ms.value)
        return out

    def set_plan(self, stage, shape=None, wg_per_cu=0):
        """Plan of one stage (name or id): shape None = runs of 256 cells,
        else (lx, ly, lz) log2 tile extents summing to 8."""
        sid = _lib.STAGES.index(stage) if isinstance(stage, str) else int(stage)
        lx, ly, lz = (-1, 0, 0) if shape is None else shape
        self._chk(self.L.pc_level_set_plan_v1(self.h, sid, lx, ly, lz, int(wg_per_cu)))

    def set_precision(self, mixed):
        """False: FP64 step (default); True: FP32 reconstruction, limiters,
        Riemann solves and intermediates, FP64 EOS inversion and update."""
        self._chk(self.L.pc_level_set_precision_v1(self.h, 1 if mixed else 0))

    def valid_tensor(self, b, out=None):
        """Box b's valid cells as a device tensor (NVAR, nz, ny, nx)."""
        import torch
        full = torch.empty(self.fab_shape(b), dtype=torch.float64, device="cuda")
        self._chk(self.L.pc_level_read_box_v1(self.h, b, C.c_void_p(full.data_ptr())))
        return full[:, 4:-4, 4:-4, 4:-4]

    def set_graph(self, steps_per_graph):
        """Replay steps from a hipGraph of `steps_per_graph` steps (0: off)."""
        self._chk(self.L.pc_level_set_graph_v1(self.h, int(steps_per_graph)))

    def graph_info(self):
        """(kernel nodes, whether the graph carries the branch DAG)."""
        n, b = C.c_int(), C.c_int()
        self._chk(self.L.pc_level_graph_info_v1(self.h, C.byref(n), C.byref(b)))
        return n.value, bool(b.value)

    def set_branches(self, on=True):
        """Run the three directions' independent stages on three streams
        (pc_level_set_branches_v1); results are unchanged."""
        self._chk(self.L.pc_level_set_branches_v1(self.h, 1 if on else 0))

    def write_plotfile(self, path, variables=None, time=0.0, step=0,
                       prob_lo=(0.0, 0.0, 0.0), stream=None):
        """AMReX plotfile directory `path` of the level (pc_level_write_plotfile_v1):
        variables = list/str of plot variable names (None: the state)."""
        if isinstance(variables, (list, tuple)):
            variables = " ".join(variables)
        v = variables.encode() if variables else None
        self._chk(self.L.pc_level_write_plotfile_v1(
            self.h, str(path).encode(), v, float(time), int(step), _d3(prob_lo),
            _stream_ptr(stream)))

    def enable_timing(self, on=True):
        self._chk(self.L.pc_level_enable_timing_v1(self.h, 1 if on else 0))

    def kernel_times(self):
        """{kernel name: average ms per timed step}"""
        n = len(_lib.KERNELS)
        a = (C.c_double * n)()
        self._chk(self.L.pc_level_kernel_times_v1(self.h, a, n))
        return dict(zip(_lib.KERNELS, list(a)))
