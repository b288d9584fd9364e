"""The plotfile format (CPU): pc_write_plotfile_v1 from host arrays against
the text of an AMReX single-level plotfile (Header, Level_0/Cell_H, FAB
headers; the layout amrex::WriteMultiLevelPlotfile writes for the
reference's PeleCAmr::writePlotFileDoit, Source/PeleCAmr.cpp:225-267),
written out here for a two-box case; the reader's round trip; the
fcompare-style comparison; and the oracle's restatement of the derived plot
variables (Source/Derive.cpp) against closed forms: the vorticity of a
solid-body rotation, the gamma-law pressure and sound speed, Mach number,
kinetic energy, speed, log density and mass fractions."""
import ctypes as C
import os

import numpy as np

import oracle_bind as ob
from particleclem_amd import plotfile as pf
from particleclem_amd.ctypes_types import a4_from_numpy, make_box

FAB_HEAD = "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"


def two_box(tmp_path):
    boxes = [([0, 0, 0], [1, 1, 1]), ([2, 0, 0], [3, 1, 1])]
    rng = np.random.default_rng(3)
    arrays = [rng.standard_normal((2, 2, 2, 2)) for _ in boxes]
    arrays[0][1] += 300.0
    arrays[1][1] += 300.0
    path = os.path.join(tmp_path, "plt00007")
    pf.write_plotfile(path, ["density", "Temp"], boxes, arrays,
                      ([0, 0, 0], [3, 1, 1]), (0.25, 0.5, 0.5), time=0.5, step=7)
    return path, boxes, arrays


def test_header_text(tmp_path):
    path, boxes, arrays = two_box(tmp_path)
    want = ("HyperCLaw-V1.1\n2\ndensity\nTemp\n3\n0.5\n0\n0 0 0 \n1 1 1 \n\n"
            "((0,0,0) (3,1,1) (0,0,0)) \n7 \n0.25 0.5 0.5 \n0\n0\n0 2 0.5\n7\n"
            "0 0.5\n0 1\n0 1\n0.5 1\n0 1\n0 1\nLevel_0/Cell\n")
    assert open(os.path.join(path, "Header")).read() == want


def test_cell_header_and_fabs(tmp_path):
    path, boxes, arrays = two_box(tmp_path)
    h0 = FAB_HEAD + "((0,0,0) (1,1,1) (0,0,0)) 2\n"
    off1 = len(h0) + 8 * 8 * 2
    mins = "".join("".join(f"{a[c].min():.15e}," for c in range(2)) + "\n" for a in arrays)
    maxs = "".join("".join(f"{a[c].max():.15e}," for c in range(2)) + "\n" for a in arrays)
    want = ("1\n1\n2\n0\n(2 0\n((0,0,0) (1,1,1) (0,0,0))\n((2,0,0) (3,1,1) (0,0,0))\n)\n"
            f"2\nFabOnDisk: Cell_D_00000 0\nFabOnDisk: Cell_D_00000 {off1}\n\n"
            f"2,2\n{mins}\n2,2\n{maxs}\n")
    assert open(os.path.join(path, "Level_0", "Cell_H")).read() == want
    raw = open(os.path.join(path, "Level_0", "Cell_D_00000"), "rb").read()
    assert raw[:len(h0)] == h0.encode()
    assert raw[off1:off1 + len(h0)] == h0.replace("((0,0,0) (1,1,1)", "((2,0,0) (3,1,1)").encode()
    body = np.frombuffer(raw[len(h0):len(h0) + 128], dtype="<f8")
    np.testing.assert_array_equal(body, arrays[0].ravel())


def test_reader_round_trip_and_compare(tmp_path):
    path, boxes, arrays = two_box(tmp_path)
    p = pf.read_plotfile(path)
    assert p.names == ["density", "Temp"] and p.step == 7 and p.time == 0.5
    assert p.dx == (0.25, 0.5, 0.5) and p.domain == ((0, 0, 0), (3, 1, 1))
    for (b, f), a in zip(zip(p.boxes, p.fabs), arrays):
        np.testing.assert_array_equal(f, a)
    d = p.dense("Temp")
    np.testing.assert_array_equal(d[:, :, :2], arrays[0][1])
    np.testing.assert_array_equal(d[:, :, 2:], arrays[1][1])
    # the same data as one box compares equal; a perturbation is measured
    one = np.concatenate([arrays[0], arrays[1]], axis=3)
    path2 = os.path.join(tmp_path, "plt_one")
    pf.write_plotfile(path2, ["density", "Temp"], [([0, 0, 0], [3, 1, 1])], [one],
                      ([0, 0, 0], [3, 1, 1]), (0.25, 0.5, 0.5))
    q = pf.read_plotfile(path2)
    assert all(v == (0.0, 0.0) for v in pf.compare(p, q).values())
    one[0, 1, 1, 3] += 1e-3
    path3 = os.path.join(tmp_path, "plt_pert")
    pf.write_plotfile(path3, ["density", "Temp"], [([0, 0, 0], [3, 1, 1])], [one],
                      ([0, 0, 0], [3, 1, 1]), (0.25, 0.5, 0.5))
    r = pf.compare(p, pf.read_plotfile(path3))
    assert abs(r["density"][0] - 1e-3) < 1e-12 and r["Temp"] == (0.0, 0.0)


def derive(kind, U, lo, hi, codes, comps, dx):
    """The oracle's plot variables of the cells [lo, hi] of U (U covers one
    ghost cell around them)."""
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    out = np.zeros((len(codes), n[2], n[1], n[0]))
    glo = [x - 1 for x in lo]
    rc = ob.lib().pc_ref_derive(kind, make_box(lo, hi), a4_from_numpy(U, glo),
                                (C.c_int * len(codes))(*codes),
                                (C.c_int * len(codes))(*comps), len(codes),
                                a4_from_numpy(out, lo), (C.c_double * 3)(*dx))
    assert rc == 0
    return out


def test_derived_variables_closed_forms():
    lo, hi = [0, 0, 0], [5, 5, 5]
    dx = (0.1, 0.2, 0.25)
    g = 8  # cells per side with one ghost
    x = (np.arange(-1, 7) + 0.5) * dx[0]
    y = (np.arange(-1, 7) + 0.5) * dx[1]
    z = (np.arange(-1, 7) + 0.5) * dx[2]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    om = 3.0
    rho = 1.0 + 0.1 * X + 0.05 * Z
    u, v, w = -om * Y, om * X, 0.2 + 0 * X
    p = 2.0 + 0.3 * Y
    gam = 1.4
    e = p / ((gam - 1.0) * rho)
    U = np.zeros((8, g, g, g))
    U[0], U[1], U[2], U[3] = rho, rho * u, rho * v, rho * w
    U[5] = rho * e
    U[4] = U[5] + 0.5 * rho * (u * u + v * v + w * w)
    U[7] = rho
    W = 28.0134
    U[6] = p * W / (rho * 8.31446261815324e7)
    codes = [0, 1, 2, 3, 4, 5, 6, 7, 8, 1]
    comps = [0, 1, 0, 0, 0, 0, 0, 0, 0, 7]
    out = derive(1, U, lo, hi, codes, comps, dx)
    V = (slice(1, -1),) * 3
    c = np.sqrt(gam * p / rho)[V]
    spd = np.sqrt(u * u + v * v + w * w)[V]
    np.testing.assert_array_equal(out[0], rho[V])
    np.testing.assert_allclose(out[1], u[V], rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(out[2], spd, rtol=1e-14)
    np.testing.assert_allclose(out[3], 0.5 * rho[V] * spd ** 2, rtol=1e-14)
    np.testing.assert_allclose(out[4], p[V], rtol=1e-14)
    np.testing.assert_allclose(out[5], c, rtol=1e-14)
    np.testing.assert_allclose(out[6], spd / c, rtol=1e-14)
    np.testing.assert_allclose(out[7], 2.0 * om, rtol=1e-12)  # |curl| of rotation
    np.testing.assert_allclose(out[8], np.log10(rho[V]), rtol=1e-15)
    np.testing.assert_allclose(out[9], 1.0, rtol=1e-15)
