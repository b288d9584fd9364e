"""Pins of the oracle's reconstructions: PPM edge values and limiter
(Source/PPM.H:28-115), the wave integrals (Source/PPM.H:131-189) and the
PLM slopes (Source/PLM.H:21-77), by properties the mathematics fixes."""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob


def ppm(s, flat=1.0):
    s = np.ascontiguousarray(s, dtype=np.float64)
    out = np.zeros(2)
    ob.lib().pc_ref_ppm_reconstruct(ob.dptr(s), flat, ob.dptr(out))
    return out


def integrals(sm, sp, sc, u, c, dtdx):
    Ip = np.zeros(3)
    Im = np.zeros(3)
    ob.lib().pc_ref_ppm_int_profile(sm, sp, sc, u, c, dtdx, ob.dptr(Ip), ob.dptr(Im))
    return Ip, Im


def plm(q, flat=1.0, order=4):
    q = np.ascontiguousarray(q, dtype=np.float64)
    return ob.lib().pc_ref_plm_slope_line(ob.dptr(q), flat, order)


def test_constant_is_preserved():
    sm, sp = ppm([2.5] * 5)
    assert sm == 2.5 and sp == 2.5


@pytest.mark.parametrize("a,b", [(1.0, 0.3), (-2.0, -0.7), (5.0, 1e-3)])
def test_linear_profile_reproduced(a, b):
    s = [a + b * i for i in range(-2, 3)]
    sm, sp = ppm(s)
    assert sm == pytest.approx(a - 0.5 * b, rel=1e-14)
    assert sp == pytest.approx(a + 0.5 * b, rel=1e-14)


def test_monotone_data_give_bounded_edges():
    rng = np.random.default_rng(5)
    for _ in range(200):
        s = np.sort(rng.normal(size=5))
        sm, sp = ppm(s)
        assert min(s[1], s[2]) <= sm <= max(s[1], s[2])
        assert min(s[2], s[3]) <= sp <= max(s[2], s[3])


def test_local_extremum_is_flattened():
    sm, sp = ppm([0.0, 1.0, 3.0, 1.0, 0.0])
    assert sm == 3.0 and sp == 3.0


def test_full_flattening_gives_constant():
    sm, sp = ppm([0.0, 0.2, 1.0, 5.0, 9.0], flat=0.0)
    assert sm == 1.0 and sp == 1.0


def test_integrals_of_a_constant_parabola():
    Ip, Im = integrals(1.5, 1.5, 1.5, 0.2, 1.0, 0.3)
    np.testing.assert_array_equal(Ip, [1.5] * 3)
    np.testing.assert_array_equal(Im, [1.5] * 3)


def test_integrals_zero_cfl_give_edges():
    Ip, Im = integrals(1.0, 2.0, 1.4, 0.3, 1.0, 0.0)
    np.testing.assert_array_equal(Ip, [2.0] * 3)
    np.testing.assert_array_equal(Im, [1.0] * 3)


def test_integral_of_linear_profile_over_wave_domain():
    # linear data a + b x on the zone [-1/2, 1/2]: the average over the last
    # sigma of the zone is a + b (1 - sigma)/2
    a, b, u, c, dtdx = 1.0, 0.4, 0.5, 0.0, 0.6
    Ip, Im = integrals(a - b / 2, a + b / 2, a, u, c, dtdx)
    sigma = u * dtdx
    assert Ip[1] == pytest.approx(a + b * (1 - sigma) / 2, rel=1e-14)


def test_plm_slopes():
    # linear data: the MC-limited slope (2nd and 4th order) is the slope
    assert plm([1, 2, 3, 4, 5], order=2) == pytest.approx(1.0)
    assert plm([1, 2, 3, 4, 5], order=4) == pytest.approx(1.0)
    # extremum -> 0; order 1 -> 0; flat = 0 -> 0
    assert plm([0, 1, 3, 1, 0]) == 0.0
    assert plm([1, 2, 3, 4, 5], order=1) == 0.0
    assert plm([1, 2, 3, 4, 5], flat=0.0) == 0.0
    # the MC limiter caps the slope at twice the one-sided difference
    assert plm([0, 0, 0, 10, 10], order=2) == pytest.approx(0.0)
    assert abs(plm([0, 0, 1, 10, 10], order=2)) <= 2.0 + 1e-15
