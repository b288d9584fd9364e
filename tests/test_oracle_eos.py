"""Pins of the oracle's EOS restatements (oracle/oracle_eos.hpp).

GammaLaw: closed forms of a calorically perfect gas.
LiDryer (ideal-gas mixture, NASA-7): NASA continuity at T_mid, the
molecular weights against the reference's PMF flame file (rho R T / (p Wbar)
= 1 on every row), the adiabatic enthalpy balances held by the reference's
PMF flame and TGReact HP-equilibrium fixtures, the e->T Newton round trip,
and the ideal-gas relations between cv, cp, gamma and the sound speed.
"""
import csv
import os

import numpy as np
import pytest

import oracle_bind as ob

HERE = os.path.dirname(os.path.abspath(__file__))
RU = 8.31446261815324e7
NAMES = ["H2", "O2", "H2O", "H", "O", "OH", "HO2", "H2O2", "N2"]
MW = np.array([2.01594, 31.9988, 18.01534, 1.00797, 15.9994, 17.00737,
               33.00677, 34.01474, 28.0134])


def hmix(T, Y):
    h = ob.eos(9, 9, T, nout=9)
    return float(np.sum(Y * h * RU * T / MW))


def cvmix(T, Y):
    return ob.eos(9, 8, T, Y=Y)


def test_gammalaw_closed_forms():
    g = 1.4
    for rho, p in [(1.0, 1.0), (0.125, 0.1), (3.7e-3, 2.5e5)]:
        e = ob.eos(1, 0, rho, p)
        assert e == pytest.approx(p / (rho * (g - 1.0)), rel=1e-15)
        c = ob.eos(1, 1, rho, p)
        assert c == pytest.approx(np.sqrt(g * p / rho), rel=1e-15)
        T = ob.eos(1, 2, rho, e, 1.0)
        assert ob.eos(1, 3, rho, T) == pytest.approx(p, rel=1e-14)
        assert ob.eos(1, 4, rho, T) == pytest.approx(c, rel=1e-14)
        assert ob.eos(1, 5, rho, T) == g
        dpde, dpdr = ob.eos(1, 10, rho, T, nout=2)
        assert dpde == pytest.approx((g - 1.0) * rho, rel=1e-15)
        assert dpdr == pytest.approx((g - 1.0) * e, rel=1e-14)


def _pmf_rows():
    with open(os.path.join(HERE, "golden", "pmf_lidryer_rows.csv")) as f:
        return list(csv.DictReader(f))


def test_molecular_weights_against_pmf_flame():
    # every row of the flame at p = 1 atm: rho R T = p Wbar within the
    # file's single-precision columns
    for r in _pmf_rows():
        X = np.maximum([float(r[n]) for n in NAMES], 0.0)
        X = X / X.sum()
        Y = X * MW / np.sum(X * MW)
        wbar = ob.eos(9, 6, Y=Y)
        assert wbar == pytest.approx(np.sum(X * MW), rel=1e-14)
        T, rho = float(r["temp"]), float(r["rho"])
        p = ob.eos(9, 3, rho, T, Y=Y)
        assert p / 1013250.0 == pytest.approx(1.0, abs=2e-5)


def test_nasa_continuity_at_tmid():
    for k in range(9):
        Y = np.zeros(9)
        Y[k] = 1.0
        lo, hi = 1000.0 * (1 - 1e-13), 1000.0
        assert cvmix(lo, Y) == pytest.approx(cvmix(hi, Y), rel=2e-6)
        assert hmix(lo, Y) == pytest.approx(hmix(hi, Y), rel=1e-6, abs=1e3)


def test_adiabatic_balance_tgreact_equilibrium():
    """h(300 K, unburnt) = h(T_ad, HP equilibrium) for the reference's
    Cantera equilibrium rows (LiDryer).  The residual is held to 1e-3 of the
    heat of formation of the water formed."""
    with open(os.path.join(HERE, "golden", "tgreact_hp_equilibrium.csv")) as f:
        rows = list(csv.DictReader(f))
    hf_water = 1.342e10  # |h_f(H2O)| per gram, erg/g
    for r in rows:
        Yu = np.zeros(9)
        Yu[0], Yu[1], Yu[8] = float(r["YH2"]), float(r["YO2"]), float(r["YN2"])
        Ye = np.array([float(r["Y%sequ" % n]) for n in NAMES])
        Tad = float(r["Tadiabatic"])
        res = hmix(Tad, Ye) - hmix(300.0, Yu)
        formed = Ye[2] - Yu[2]
        assert abs(res) <= 1e-3 * formed * hf_water + 1e-3, (Tad, res, formed)


def test_adiabatic_balance_pmf_flame():
    rows = _pmf_rows()

    def h_of(r):
        X = np.maximum([float(r[n]) for n in NAMES], 0.0)
        X = X / X.sum()
        Y = X * MW / np.sum(X * MW)
        return hmix(float(r["temp"]), Y), Y
    hu, Yu = h_of(rows[0])
    hb, Yb = h_of(rows[-1])
    Tu, Tb = float(rows[0]["temp"]), float(rows[-1]["temp"])
    cp = cvmix(Tb, Yb) + RU / ob.eos(9, 6, Y=Yb)
    assert abs(hb - hu) <= 1e-4 * cp * (Tb - Tu)


@pytest.mark.parametrize("T", [250.0, 300.0, 700.0, 999.5, 1000.5, 1500.0, 2500.0])
def test_newton_e_to_T_round_trip(T):
    rng = np.random.default_rng(7)
    for _ in range(5):
        Y = rng.random(9)
        Y /= Y.sum()
        e = ob.eos(9, 0, 1.0e-3, ob.eos(9, 3, 1.0e-3, T, Y=Y), Y=Y)
        for guess in (T * 0.7, T * 1.3, 1.0):
            Tn = ob.eos(9, 2, 1.0e-3, e, guess, Y=Y)
            assert Tn == pytest.approx(T, rel=1e-12)


def test_ideal_gas_relations_mixture():
    rng = np.random.default_rng(3)
    for _ in range(10):
        Y = rng.random(9)
        Y /= Y.sum()
        T = rng.uniform(300, 2500)
        rho = rng.uniform(1e-4, 1e-2)
        wbar = ob.eos(9, 6, Y=Y)
        assert wbar == pytest.approx(1.0 / np.sum(Y / MW), rel=1e-14)
        p = ob.eos(9, 3, rho, T, Y=Y)
        assert p == pytest.approx(rho * RU * T / wbar, rel=1e-14)
        cv = cvmix(T, Y)
        # cv from a centred difference of e(T)
        dT = 1e-3
        e1 = ob.eos(9, 0, rho, ob.eos(9, 3, rho, T + dT, Y=Y), Y=Y)
        e0 = ob.eos(9, 0, rho, ob.eos(9, 3, rho, T - dT, Y=Y), Y=Y)
        if abs(T - 1000.0) > 1.0:
            assert (e1 - e0) / (2 * dT) == pytest.approx(cv, rel=1e-6)
        gam = (cv + RU / wbar) / cv
        assert ob.eos(9, 5, rho, T, Y=Y) == pytest.approx(gam, rel=1e-14)
        c = ob.eos(9, 1, rho, p, Y=Y)
        assert c == pytest.approx(np.sqrt(gam * p / rho), rel=1e-14)
        assert ob.eos(9, 4, rho, T, Y=Y) == pytest.approx(c, rel=1e-13)


def test_horner_and_mixture_forms_agree_with_term_by_term():
    """The Horner/FMA species energies against h/RT summed term by term
    (e_k = (h_k/RT - 1) R T / W_k), and the mixture-coefficient cv against
    the species-sum cv, on both sides of T_mid."""
    rng = np.random.default_rng(11)
    for T in (200.0, 650.0, 999.0, 1001.0, 1700.0, 3200.0):
        ek = ob.eos(9, 7, T, nout=9)
        h = ob.eos(9, 9, T, nout=9)
        ref = (h - 1.0) * RU * T / MW
        scale = np.abs(h) * RU * T / MW
        assert np.all(np.abs(ek - ref) <= 1e-14 * scale), (T, ek - ref)
        for _ in range(4):
            Y = rng.random(9)
            Y /= Y.sum()
            cs = ob.eos(9, 8, T, Y=Y)
            cm = ob.eos(9, 11, T, Y=Y)
            assert cm == pytest.approx(cs, rel=1e-14)
