"""ctypes mirrors of the POD structs in include/pc_hydro.h.

`pc_array4_t` is the field set of amrex::Array4<double>; `pc_box_t` is a
cell-centred amrex::Box.  Helpers build views of numpy arrays or raw device
pointers laid out as AMReX FABs (i fastest, component slowest).
"""
import ctypes as C

import numpy as np


class Array4(C.Structure):
    _fields_ = [
        ("p", C.c_void_p),
        ("jstride", C.c_longlong),
        ("kstride", C.c_longlong),
        ("nstride", C.c_longlong),
        ("begin", C.c_int * 3),
        ("end", C.c_int * 3),
        ("ncomp", C.c_int),
    ]


class Box(C.Structure):
    _fields_ = [("lo", C.c_int * 3), ("hi", C.c_int * 3)]


class Params(C.Structure):
    _fields_ = [
        ("ppm_type", C.c_int),
        ("plm_iorder", C.c_int),
        ("use_flattening", C.c_int),
        ("use_hybrid_weno", C.c_int),
        ("weno_scheme", C.c_int),
        ("difmag", C.c_double),
        ("cfl", C.c_double),
        ("init_shrink", C.c_double),
        ("change_max", C.c_double),
        ("max_dt", C.c_double),
        ("fixed_dt", C.c_double),
        ("dt_cutoff", C.c_double),
        ("allow_small_energy", C.c_int),
        ("allow_negative_energy", C.c_int),
        ("dual_energy_update_E_from_e", C.c_int),
        ("dual_energy_eta2", C.c_double),
        ("do_mol", C.c_int),
        ("mol_iorder", C.c_int),
        ("use_laxf_flux", C.c_int),
        ("stop_time", C.c_double),
    ]


class LevelDesc(C.Structure):
    _fields_ = [
        ("nboxes", C.c_int),
        ("boxes", C.POINTER(Box)),
        ("domlo", C.c_int * 3),
        ("domhi", C.c_int * 3),
        ("bclo", C.c_int * 3),
        ("bchi", C.c_int * 3),
        ("periodic_self", C.c_int),
        ("dx", C.c_double * 3),
        ("problo", C.c_double * 3),
    ]


# reference defaults (Source/Params/_cpp_parameters)
DEFAULTS = dict(
    ppm_type=0, plm_iorder=4, use_flattening=1, use_hybrid_weno=0,
    weno_scheme=1, difmag=0.1, cfl=0.8, init_shrink=1.0, change_max=1.1,
    max_dt=1.0e200, fixed_dt=-1.0, dt_cutoff=0.0, allow_small_energy=1,
    allow_negative_energy=1, dual_energy_update_E_from_e=1,
    dual_energy_eta2=1.0e-4, do_mol=0, mol_iorder=2, use_laxf_flux=0,
    stop_time=-1.0,
)


def make_params(**kw):
    d = dict(DEFAULTS)
    for k, v in kw.items():
        if k not in d:
            raise KeyError(f"unknown parameter {k}")
        d[k] = v
    return Params(**d)


def make_box(lo, hi):
    return Box((C.c_int * 3)(*lo), (C.c_int * 3)(*hi))


def grow(lo, hi, n):
    return [x - n for x in lo], [x + n for x in hi]


def box_shape(lo, hi):
    """numpy shape (k, j, i) of an inclusive box."""
    return (hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)


def a4_from_ptr(ptr, lo, hi, ncomp):
    """View of a contiguous FAB over the inclusive box [lo, hi]."""
    nx = hi[0] - lo[0] + 1
    ny = hi[1] - lo[1] + 1
    nz = hi[2] - lo[2] + 1
    a = Array4()
    a.p = ptr
    a.jstride = nx
    a.kstride = nx * ny
    a.nstride = nx * ny * nz
    a.begin = (C.c_int * 3)(*lo)
    a.end = (C.c_int * 3)(hi[0] + 1, hi[1] + 1, hi[2] + 1)
    a.ncomp = ncomp
    return a


def a4_from_numpy(arr, lo):
    """arr: float64 C-contiguous array of shape (ncomp, nz, ny, nx) whose
    element [n, k, j, i] is cell (lo+i, lo+j, lo+k) comp n."""
    assert arr.dtype == np.float64 and arr.flags["C_CONTIGUOUS"]
    nc, nz, ny, nx = arr.shape
    hi = [lo[0] + nx - 1, lo[1] + ny - 1, lo[2] + nz - 1]
    return a4_from_ptr(arr.ctypes.data, lo, hi, nc)


def null_a4():
    return Array4()
