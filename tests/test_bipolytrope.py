"""CPU checks of the served workload's bipolytrope profile (host code of
libbh_potential.so) against what the mathematics fixes: the n = 1 closed
form, published Lane-Emden surface radii and masses, and the interface
placement of the composite star."""
import math

import numpy as np
import pytest

from bipoly_hydro_amd import potential as P

# first zero xi_1 and -xi_1^2 theta'(xi_1) of the Lane-Emden functions
# (Chandrasekhar 1939, table 4)
LE = {1.5: (3.65375, 2.71406), 3.0: (6.89685, 2.01824)}


def mass_integral(t):
    u = np.linspace(0.0, 1.0, len(t))
    return np.trapezoid(u * u * t, u)


@pytest.mark.parametrize("core_frac", [0.0, 0.3, 0.7])
def test_n1_composite_is_the_closed_form(core_frac):
    t, R, xi = P.bipolytrope_profile(1.0, 1.0, core_frac)
    u = np.linspace(0.0, 1.0, len(t))
    ref = np.where(u > 0, np.sin(math.pi * u) / np.where(u > 0, math.pi * u, 1.0), 1.0)
    assert abs(R - math.pi) < 1e-8
    assert np.abs(t - ref).max() < 1e-7


@pytest.mark.parametrize("n", [1.5, 3.0])
def test_single_polytropes_match_published_constants(n):
    xi1, mass = LE[n]
    for args in ((n, n, 0.0), (n, n, 1.0)):
        t, R, _ = P.bipolytrope_profile(*args)
        assert abs(R - xi1) < 2e-5
        # M = 4 pi xi1^3 int u^2 f(u) du = 4 pi (-xi1^2 theta'(xi1))
        assert abs(R ** 3 * mass_integral(t) - mass) < 2e-4


def test_core_envelope_composite():
    t, R, xi_i = P.bipolytrope_profile(3.0, 1.5, 0.3)
    assert abs(xi_i / R - 0.3) < 1e-6
    assert t[0] == 1.0 and t[-1] == 0.0
    assert np.all(np.diff(t) <= 1e-15)  # density falls monotonically
    k = int(round(0.3 * (len(t) - 1)))
    assert abs(t[k + 1] - t[k - 1]) < 5e-3  # no jump at the interface


def test_bad_profile_arguments_are_errors():
    with pytest.raises(P.BHError):
        P.bipolytrope_profile(6.0, 1.5, 0.3)
    with pytest.raises(P.BHError):
        P.bipolytrope_profile(3.0, 1.5, 1.5)
