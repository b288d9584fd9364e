"""Read and compare AMReX plotfiles (single level, cell-centred).

read_plotfile(path) parses what pc_level_write_plotfile_v1 writes, and any
plotfile of the same format (amrex::WriteMultiLevelPlotfile output, e.g. a
PeleC run's plt* directory; level 0 only): the Header, the VisMF header
Level_0/Cell_H and the FABs it points at.

compare(a, b) reports, per variable, the largest absolute difference and
that difference relative to max |a| over the level, as AMReX's fcompare
tool does with its default norm (the reference's regression tests compare
plotfiles that way, Tests/CMakeLists.txt:89-96).
"""
import os
import re

import numpy as np

_BOX = re.compile(r"\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\) "
                  r"\((\d+),(\d+),(\d+)\)\)")


class Plotfile:
    def __init__(self, names, time, step, domain, prob_lo, prob_hi, dx, boxes,
                 fabs, header_lines):
        self.names = names
        self.time = time
        self.step = step
        self.domain = domain          # ((lo), (hi))
        self.prob_lo = prob_lo
        self.prob_hi = prob_hi
        self.dx = dx
        self.boxes = boxes            # [((lo), (hi)), ...]
        self.fabs = fabs              # [array (nvars, nz, ny, nx), ...]
        self.header_lines = header_lines

    def var(self, name):
        """[(box, array (nz, ny, nx))] of one variable."""
        k = self.names.index(name)
        return [(b, f[k]) for b, f in zip(self.boxes, self.fabs)]

    def dense(self, name):
        """The variable on the whole domain as one array (nz, ny, nx)."""
        (lo, hi) = self.domain
        out = np.full(tuple(hi[d] - lo[d] + 1 for d in (2, 1, 0)), np.nan)
        for (blo, bhi), a in self.var(name):
            out[blo[2] - lo[2]:bhi[2] - lo[2] + 1, blo[1] - lo[1]:bhi[1] - lo[1] + 1,
                blo[0] - lo[0]:bhi[0] - lo[0] + 1] = a
        return out


def _box(text):
    m = _BOX.search(text)
    if not m:
        raise ValueError(f"not a box: {text!r}")
    v = [int(x) for x in m.groups()]
    return (tuple(v[0:3]), tuple(v[3:6]))


def read_plotfile(path):
    with open(os.path.join(path, "Header")) as f:
        lines = f.read().split("\n")
    it = iter(lines)
    version = next(it)
    if not version.startswith("HyperCLaw"):
        raise ValueError(f"unknown plotfile version {version!r}")
    nvar = int(next(it))
    names = [next(it) for _ in range(nvar)]
    dim = int(next(it))
    if dim != 3:
        raise ValueError("only 3D plotfiles are supported")
    time = float(next(it))
    finest = int(next(it))
    prob_lo = tuple(float(x) for x in next(it).split())
    prob_hi = tuple(float(x) for x in next(it).split())
    next(it)  # refinement ratios
    domain = _box(next(it))
    step = int(next(it).split()[0])
    dx = tuple(float(x) for x in next(it).split())
    for _ in range(finest):
        next(it)
    next(it)  # coordinate system
    next(it)  # boundary width
    lev, ngrid, _t = next(it).split()
    ngrid = int(ngrid)
    next(it)  # level steps
    for _ in range(ngrid * dim):
        next(it)
    prefix = next(it)
    base = os.path.join(path, prefix)
    with open(base + "_H") as f:
        vh = f.read().split("\n")
    ncomp = int(vh[2])
    nb = int(vh[4].strip("(").split()[0])
    boxes = [_box(vh[5 + i]) for i in range(nb)]
    k = 5 + nb + 1
    nfod = int(vh[k])
    fods = []
    for i in range(nfod):
        tag, fname, off = vh[k + 1 + i].split()
        fods.append((fname, int(off)))
    fabs = []
    for (blo, bhi), (fname, off) in zip(boxes, fods):
        with open(os.path.join(os.path.dirname(base), fname), "rb") as f:
            f.seek(off)
            head = f.readline().decode()
            if not head.startswith("FAB "):
                raise ValueError(f"bad FAB header {head!r}")
            lo, hi = _box(head[head.rfind("(("):])
            nc = int(head.split()[-1])
            little = "(8, (8 7 6 5 4 3 2 1))" in head
            shape = (nc, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
            a = np.fromfile(f, dtype="<f8" if little else ">f8",
                            count=int(np.prod(shape)))
            fabs.append(a.reshape(shape))
        if nc != ncomp:
            raise ValueError("FAB component count disagrees with Cell_H")
    return Plotfile(names, time, step, domain, prob_lo, prob_hi, dx, boxes, fabs,
                    lines)


def compare(a, b):
    """{variable: (max |a-b|, max |a-b| / max |a|)} over the common variables
    of two plotfiles of the same domain (any box decomposition)."""
    out = {}
    for name in a.names:
        if name not in b.names:
            continue
        x, y = a.dense(name), b.dense(name)
        d = float(np.nanmax(np.abs(x - y)))
        s = float(np.nanmax(np.abs(x)))
        out[name] = (d, d / s if s > 0 else d)
    return out


def write_plotfile(path, names, boxes, arrays, domain, dx,
                   prob_lo=(0.0, 0.0, 0.0), time=0.0, step=0):
    """Write host arrays (one (nvars, nz, ny, nx) float64 array per box
    (lo, hi)) as an AMReX plotfile through pc_write_plotfile_v1."""
    import ctypes as C

    from . import _lib
    from .ctypes_types import Box, make_box
    L = _lib.lib()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
    for a, (lo, hi) in zip(arrs, boxes):
        assert a.shape == (len(names),) + tuple(hi[d] - lo[d] + 1 for d in (2, 1, 0))
    nm = (C.c_char_p * len(names))(*[n.encode() for n in names])
    bx = (Box * len(boxes))(*[make_box(lo, hi) for lo, hi in boxes])
    dp = (C.POINTER(C.c_double) * len(arrs))(
        *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])
    i3 = lambda v: (C.c_int * 3)(*v)
    d3 = lambda v: (C.c_double * 3)(*v)
    _lib.check(L.pc_write_plotfile_v1(str(path).encode(), len(names), nm, len(boxes),
                                      bx, dp, i3(domain[0]), i3(domain[1]),
                                      d3(prob_lo), d3(dx), float(time), int(step)))
