"""Pins of the oracle's method-of-lines path (CPU): the Lax-Friedrichs flux
against closed forms, the MOL step against the invariants of the exact
algorithm (uniform state, discrete conservation, species/density
consistency, axis permutation, mirror and box decomposition), its
convergence to the reference's exact Sod solution (the sod-3 configuration
of Exec/RegTests/Sod: MOL, CFL 0.3, slip walls at the tube ends) and the
order of accuracy on a smooth wave: second order with slopes
(plm_iorder != 1), first order without — a dropped or mis-signed slope
term breaks the former."""
import ctypes as C
import csv
import os

import numpy as np
import pytest

import oracle_bind as ob
from particleclem_amd import workloads as wl
from parity_util import VALID, oracle_level, run_oracle

HERE = os.path.dirname(os.path.abspath(__file__))
MOL = dict(do_mol=1)
LAXF = dict(do_mol=1, use_laxf_flux=1)

GAMMA = 1.4


def _gamma_state(rho, u, v, w, p):
    e = p / ((GAMMA - 1.0) * rho)
    c = np.sqrt(GAMMA * p / rho)
    return e, c


@pytest.mark.parametrize("u", [0.7, -1.3, 0.0])
def test_laxf_identical_states_give_physical_flux(u):
    rho, v, w, p = 1.3, 0.4, -0.2, 2.1
    f = ob.laxf(1, [rho, u, v, w, p], [1.0], [rho, u, v, w, p], [1.0])
    e, c = _gamma_state(rho, u, v, w, p)
    E = e + 0.5 * (u * u + v * v + w * w)
    np.testing.assert_allclose(f["ustar"], u, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(f["maxeig"], abs(u) + c, rtol=1e-15)
    np.testing.assert_allclose(
        [f["f_rho"], f["f_un"], f["f_ut"], f["f_utt"], f["f_eden"], f["f_eint"]],
        [rho * u, rho * u * u + p, rho * v * u, rho * w * u,
         (rho * E + p) * u, rho * e * u], rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(f["f_rhoY"], [rho * u], rtol=1e-14, atol=1e-14)


def test_laxf_identical_mixture_states_nine_species():
    Y = np.array([0.02, 0.2, 0.001, 0.002, 0.05, 0.003, 0.0005, 0.0002, 0.0])
    Y[-1] = 1.0 - Y[:-1].sum()
    rho, u, v, w, p = 1.1e-3, 250.0, -30.0, 12.0, 1.2e6
    f = ob.laxf(9, [rho, u, v, w, p], Y, [rho, u, v, w, p], Y)
    e = ob.eos(9, 0, rho, p, Y=Y)  # RYP2E
    E = e + 0.5 * (u * u + v * v + w * w)
    np.testing.assert_allclose(f["f_rhoY"], rho * u * Y, rtol=1e-13, atol=1e-20)
    np.testing.assert_allclose(f["f_rho"], rho * u, rtol=1e-13)
    np.testing.assert_allclose(f["f_eden"], (rho * E + p) * u, rtol=1e-13)
    np.testing.assert_allclose(f["f_un"], rho * u * u + p, rtol=1e-13)


def test_laxf_contact_at_rest_is_pure_dissipation():
    """u = 0 and equal p on both sides: ustar = 0 exactly, the central part
    vanishes and the flux is -0.5 * 0.5 (cl + cr) * (U_R - U_L)."""
    rl, rr, p = 1.0, 0.25, 0.6
    f = ob.laxf(1, [rl, 0.0, 0.0, 0.0, p], [1.0], [rr, 0.0, 0.0, 0.0, p], [1.0])
    _, cl = _gamma_state(rl, 0, 0, 0, p)
    _, cr = _gamma_state(rr, 0, 0, 0, p)
    a = 0.5 * (cl + cr)
    assert f["ustar"] == 0.0
    np.testing.assert_allclose(f["maxeig"], a, rtol=1e-15)
    np.testing.assert_allclose(f["f_rho"], -0.5 * a * (rr - rl), rtol=1e-15)
    np.testing.assert_allclose(f["f_un"], p, rtol=1e-15)  # 0.5 (p + p)
    el = p / ((GAMMA - 1) * rl)
    er = p / ((GAMMA - 1) * rr)
    np.testing.assert_allclose(f["f_eden"], -0.5 * a * (rr * er - rl * el),
                               rtol=1e-14, atol=1e-15)


def test_laxf_mirror_symmetry():
    L = [1.0, 0.3, 0.1, -0.4, 1.0]
    R = [0.3, -0.5, 0.2, 0.6, 0.4]
    a = ob.laxf(1, L, [1.0], R, [1.0])
    Lm = [R[0], -R[1], R[2], R[3], R[4]]
    Rm = [L[0], -L[1], L[2], L[3], L[4]]
    b = ob.laxf(1, Lm, [1.0], Rm, [1.0])
    np.testing.assert_allclose(b["f_rho"], -a["f_rho"], rtol=1e-14)
    np.testing.assert_allclose(b["f_un"], a["f_un"], rtol=1e-14)
    np.testing.assert_allclose(b["f_ut"], -a["f_ut"], rtol=1e-14)
    np.testing.assert_allclose(b["f_eden"], -a["f_eden"], rtol=1e-14)


def periodic(w):
    w.bclo = [0, 0, 0]
    w.bchi = [0, 0, 0]
    return w


@pytest.mark.parametrize("over", [MOL, LAXF], ids=["riemann", "laxf"])
def test_uniform_state_is_preserved(over):
    w = periodic(wl.sod(n=12, pert=0.0, nsteps=0))
    for P in w.prims:
        P[0] = 0.7
        P[1] = 0.4
        P[2], P[3], P[4] = 0.3, -0.2, 0.1
    r, _ = run_oracle(w, 3, over)
    V = r[0][VALID]
    for n in range(V.shape[0]):
        assert np.ptp(V[n]) <= 1e-15 * max(1.0, np.abs(V[n]).max()), n


def test_mol_source_of_uniform_state_is_zero():
    w = periodic(wl.premixed(n=8, box=8, nsteps=0, sigma=0.0))
    H, prm = oracle_level(w, MOL)
    for P in w.prims:
        P[...] = P[:, :1, :1, :1]
    H.set_prims(w)
    H.fill_ghosts()
    lo, hi = w.boxes[0]
    src = ob.mol_src_box(9, H.fab(0), lo, hi, w.dx, prm)
    assert np.all(src == 0.0)


@pytest.mark.parametrize("over", [MOL, LAXF], ids=["riemann", "laxf"])
@pytest.mark.parametrize("kind", ["sod", "premixed"])
def test_conservation_periodic(kind, over):
    if kind == "sod":
        w = periodic(wl.sod(n=16, nsteps=0))
        comps = [0, 1, 2, 3, 4, 7]
    else:
        w = periodic(wl.premixed(n=16, box=8, nsteps=0))
        comps = [0, 1, 2, 3, 4] + list(range(7, 16))
    H, prm = oracle_level(w, over)

    def sums():
        return np.array([sum(H.valid(b)[n].sum() for b in range(len(w.boxes)))
                         for n in comps])
    s0 = sums()
    dt = H.estdt(prm)
    for _ in range(3):
        dt = H.step(prm, 1, dt)
        s = sums()
        scale = np.array([sum(np.abs(H.valid(b)[n]).sum() for b in range(len(w.boxes)))
                          for n in comps])
        # momentum sums start near zero: measure them against the momentum
        # scale sqrt(rho rhoE) of the state (rho c up to a constant)
        pscale = sum(np.sqrt(np.abs(H.valid(b)[0] * H.valid(b)[4])).sum()
                     for b in range(len(w.boxes)))
        for k, n in enumerate(comps):
            sc = max(scale[k], pscale) if n in (1, 2, 3) else scale[k]
            assert abs(s[k] - s0[k]) <= 1e-13 * sc, (n, s[k], s0[k])


def test_species_sum_equals_density():
    w = wl.premixed(n=16, box=8, nsteps=0)
    r, _ = run_oracle(w, 3, MOL)
    for f in r:
        V = f[VALID]
        np.testing.assert_allclose(V[7:].sum(axis=0), V[0], rtol=2e-15)


@pytest.mark.parametrize("over", [MOL, LAXF], ids=["riemann", "laxf"])
def test_axis_permutation_is_exact(over):
    outs = []
    for idir in range(3):
        w = wl.sod(n=16, ny=6, nz=6, idir=idir, pert=0.0, nsteps=0)
        r, _ = run_oracle(w, 3, over)
        V = r[0][VALID]
        if idir == 1:
            V = np.transpose(V, (0, 1, 3, 2))[[0, 2, 1, 3, 4, 5, 6, 7]]
        elif idir == 2:
            V = np.transpose(V, (0, 3, 2, 1))[[0, 3, 2, 1, 4, 5, 6, 7]]
        outs.append(V)
    for V in outs[1:]:
        np.testing.assert_array_equal(V[:, 0, 0, :], outs[0][:, 0, 0, :])


def test_mirror_symmetry():
    a, _ = run_oracle(wl.sod(n=16, ny=4, nz=4, pert=0.0), 3, MOL)
    b, _ = run_oracle(wl.sod(n=16, ny=4, nz=4, pert=0.0, mirror=True), 3, MOL)
    A, B = a[0][VALID], b[0][VALID][:, :, :, ::-1].copy()
    B[1] = -B[1]
    np.testing.assert_allclose(B, A, rtol=1e-13, atol=1e-15)


def test_box_decomposition_is_exact():
    one, _ = run_oracle(wl.sod(n=32, ny=16, nz=16, nsteps=0), 2, MOL)
    many, _ = run_oracle(wl.sod(n=32, ny=16, nz=16, max_grid=8, nsteps=0), 2, MOL)
    w = wl.sod(n=32, ny=16, nz=16, max_grid=8, nsteps=0)
    big = one[0][VALID]
    for b, (lo, hi) in enumerate(w.boxes):
        sub = big[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
        np.testing.assert_array_equal(many[b][VALID], sub)


def _advance_to(w, t_end, over):
    H, prm = oracle_level(w, over)
    t, dt = 0.0, H.estdt(prm)
    while t < t_end - 1e-14:
        dt = min(dt, t_end - t)
        nxt = H.step(prm, 1, dt)
        t += dt
        dt = nxt
    return H


def _exact():
    with open(os.path.join(HERE, "golden", "sod_exact_t0p1.csv")) as f:
        rows = list(csv.DictReader(f))
    x = np.array([float(r["x"]) for r in rows])
    rho = np.array([float(r["density"]) for r in rows])
    return x, rho


@pytest.mark.parametrize("over,bound", [(MOL, 0.01), (LAXF, 0.015)],
                         ids=["riemann", "laxf"])
def test_sod_converges_to_exact_solution(over, bound):
    x_ex, rho_ex = _exact()
    errs = []
    for n in (32, 64, 128):
        w = wl.sod(n=n, ny=4, nz=4, pert=0.0, nsteps=0, xbc="slipwall")
        H = _advance_to(w, 0.1, over)
        rho = H.valid(0)[0, 0, 0, :]
        x = (np.arange(n) + 0.5) / n - 0.5
        m = (x > x_ex.min()) & (x < x_ex.max())
        errs.append(np.mean(np.abs(rho[m] - np.interp(x[m], x_ex, rho_ex))))
    print(over, errs)
    assert errs[2] < bound
    assert errs[0] / errs[1] > 1.4 and errs[1] / errs[2] > 1.4, errs


def _wave_error(n, plm_iorder, t_end=0.25):
    w = wl.density_wave(n=n, ny=4, nz=4, cfl=0.3)
    H = _advance_to(w, t_end, dict(do_mol=1, plm_iorder=plm_iorder))
    rho = H.valid(0)[0, 0, 0, :]
    x = (np.arange(n) + 0.5) / n
    exact = 1.0 + 0.2 * np.sin(2.0 * np.pi * (x - t_end))
    return np.mean(np.abs(rho - exact))


def test_order_of_accuracy_on_a_smooth_wave():
    e2 = [_wave_error(n, 2) for n in (16, 32, 64)]
    e1 = [_wave_error(n, 1) for n in (16, 32, 64)]
    r2 = np.log2(e2[1] / e2[2])
    r1 = np.log2(e1[1] / e1[2])
    print("slopes", e2, r2, "no slopes", e1, r1)
    assert r2 > 1.6, e2
    assert 0.7 < r1 < 1.3, e1
    assert e2[2] < 0.2 * e1[2]


def test_mol_iorder_is_validated():
    w = wl.sod(n=8, ny=4, nz=4, nsteps=0)
    H, prm = oracle_level(w, dict(do_mol=1, mol_iorder=3))
    d = C.c_double(1e-3)
    rc = ob.lib().pc_ref_level_step(1, C.byref(H.desc), C.byref(prm),
                                    ob.dptr(H.arena), 1, C.byref(d), None, 0)
    assert rc == -10
