"""Whole-step pins of the oracle (CPU): invariants that hold for the exact
algorithm whatever the data — uniform-state preservation, discrete
conservation on a periodic domain (the Tests/test_masscons.py pattern, rtol
1e-13), species/density consistency, axis-permutation, mirror and
box-decomposition invariance — and convergence to the reference's exact Sod
solution (Exec/RegTests/EB-C7/exact_soln, t = 0.1)."""
import csv
import os

import numpy as np
import pytest

import oracle_bind as ob
from particleclem_amd import workloads as wl
from parity_util import VALID, oracle_level, run_oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def periodic(w):
    w.bclo = [0, 0, 0]
    w.bchi = [0, 0, 0]
    return w


def test_uniform_state_is_preserved():
    w = wl.sod(n=12, pert=0.0, nsteps=0)
    for P in w.prims:
        P[0] = 0.7
        P[1] = 0.4
        P[2], P[3], P[4] = 0.3, -0.2, 0.1
    periodic(w)
    r, _ = run_oracle(w, 3)
    V = r[0][VALID]
    for n in range(V.shape[0]):
        assert np.ptp(V[n]) <= 1e-15 * max(1.0, np.abs(V[n]).max()), n


@pytest.mark.parametrize("kind", ["sod", "premixed"])
def test_conservation_periodic(kind):
    if kind == "sod":
        w = periodic(wl.sod(n=16, nsteps=0))
        comps = [0, 1, 2, 3, 4, 7]
    else:
        w = periodic(wl.premixed(n=16, box=8, nsteps=0))
        comps = [0, 1, 2, 3, 4] + list(range(7, 16))
    H, prm = oracle_level(w)

    def sums():
        return np.array([sum(H.valid(b)[n].sum() for b in range(len(w.boxes)))
                         for n in comps])
    s0 = sums()
    dt = H.estdt(prm)
    for _ in range(4):
        dt = H.step(prm, 1, dt)
        s = sums()
        scale = np.array([sum(np.abs(H.valid(b)[n]).sum() for b in range(len(w.boxes)))
                          for n in comps])
        for k, n in enumerate(comps):
            assert abs(s[k] - s0[k]) <= 1e-13 * scale[k], (n, s[k], s0[k])


def test_species_sum_equals_density():
    w = wl.premixed(n=16, box=8, nsteps=0)
    r, _ = run_oracle(w, 3)
    for f in r:
        V = f[VALID]
        np.testing.assert_allclose(V[7:].sum(axis=0), V[0], rtol=2e-15)


def test_axis_permutation_is_exact():
    outs = []
    for idir in range(3):
        w = wl.sod(n=16, ny=6, nz=6, idir=idir, pert=0.0, nsteps=0)
        r, _ = run_oracle(w, 3)
        V = r[0][VALID]               # (nvar, z, y, x)
        # bring the tube axis to x
        if idir == 1:
            V = np.transpose(V, (0, 1, 3, 2))[[0, 2, 1, 3, 4, 5, 6, 7]]
        elif idir == 2:
            V = np.transpose(V, (0, 3, 2, 1))[[0, 3, 2, 1, 4, 5, 6, 7]]
        outs.append(V)
    # rows sorted by the permuted axis order: compare the tube profile
    for V in outs[1:]:
        np.testing.assert_array_equal(V[:, 0, 0, :], outs[0][:, 0, 0, :])


def test_mirror_symmetry():
    a, _ = run_oracle(wl.sod(n=16, ny=4, nz=4, pert=0.0), 3)
    b, _ = run_oracle(wl.sod(n=16, ny=4, nz=4, pert=0.0, mirror=True), 3)
    A, B = a[0][VALID], b[0][VALID][:, :, :, ::-1].copy()
    B[1] = -B[1]
    np.testing.assert_allclose(B, A, rtol=1e-13, atol=1e-15)


def test_box_decomposition_is_exact():
    one, _ = run_oracle(wl.sod(n=32, ny=16, nz=16, nsteps=0), 2)
    many, _ = run_oracle(wl.sod(n=32, ny=16, nz=16, max_grid=8, nsteps=0), 2)
    w = wl.sod(n=32, ny=16, nz=16, max_grid=8, nsteps=0)
    big = one[0][VALID]
    for b, (lo, hi) in enumerate(w.boxes):
        sub = big[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
        np.testing.assert_array_equal(many[b][VALID], sub)


def _exact():
    with open(os.path.join(HERE, "golden", "sod_exact_t0p1.csv")) as f:
        rows = list(csv.DictReader(f))
    x = np.array([float(r["x"]) for r in rows])
    rho = np.array([float(r["density"]) for r in rows])
    return x, rho


def test_sod_converges_to_exact_solution():
    x_ex, rho_ex = _exact()
    errs = []
    for n in (32, 64, 128):
        w = wl.sod(n=n, ny=4, nz=4, pert=0.0, nsteps=0)
        H, prm = oracle_level(w)
        t, dt = 0.0, H.estdt(prm)
        while t < 0.1 - 1e-14:
            dt = min(dt, 0.1 - t)
            nxt = H.step(prm, 1, dt)
            t += dt
            dt = nxt
        rho = H.valid(0)[0, 0, 0, :]
        x = (np.arange(n) + 0.5) / n - 0.5
        m = (x > x_ex.min()) & (x < x_ex.max())
        errs.append(np.mean(np.abs(rho[m] - np.interp(x[m], x_ex, rho_ex))))
    assert errs[2] < 0.01
    assert errs[0] / errs[1] > 1.5 and errs[1] / errs[2] > 1.5, errs
