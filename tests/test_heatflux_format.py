"""The heatflux2 record layout (write_heat_flux2, heatcurrent.f90:730-740) and the Fortran
ES edit descriptor it uses, against values whose Fortran output is fixed by the standard."""
import importlib

import pytest


@pytest.fixture(scope="module")
def md():
    return importlib.import_module("dlpoly2-mbpol-heat_amd.md")


@pytest.mark.parametrize("v,w,d,out", [
    (1.0e-3, 12, 5, " 1.00000E-03"),
    (0.2, 12, 5, " 2.00000E-01"),
    (-123.456, 15, 7, " -1.2345600E+02"),
    (0.0, 15, 7, "  0.0000000E+00"),
    (9.99999996, 15, 7, "  1.0000000E+01"),      # rounds up into the next decade
    (1.0e100, 15, 7, "  1.0000000+100"),         # three-digit exponent drops the letter
    (-2.5e-120, 15, 7, " -2.5000000-120"),
])
def test_fortran_es(md, v, w, d, out):
    assert md.fortran_es(v, w, d) == out


def test_record_layout(md):
    rec = md.heatflux2_record(0.0002, [1e-3, -2e-3, 3e-3], [0.5, 0.25, -0.125], [1.0, 2.0, 3.0])
    lines = rec.splitlines()
    assert len(lines) == 4
    assert lines[0] == " 2.00000E-04"
    assert lines[1] == "  1.0000000E+00 -2.0000000E+00  3.0000000E+00"     # x junit = 1e3
    assert [len(l) for l in lines[1:]] == [45, 45, 45]
    assert lines[3] == "  1.0000000E+03  2.0000000E+03  3.0000000E+03"


def test_check_record_layout(md):
    """write_check's CHECK_ENERGY_STRESS record (heat_check.f90:33-35): es12.5, 2es15.7, 9es15.7."""
    rec = md.check_record(0.0004, 418.4 * 2.0, -418.4 * 30.5, [[1.5, 0, -2], [0, 3, 0], [-2, 0, 4.25]], strsunit=2.0)
    lines = rec.splitlines()
    assert lines[0] == " 4.00000E-04"
    assert lines[1] == "  2.0000000E+00 -3.0500000E+01"
    assert lines[2].split() == ["3.0000000E+00", "0.0000000E+00", "-4.0000000E+00", "0.0000000E+00",
                                "6.0000000E+00", "0.0000000E+00", "-4.0000000E+00", "0.0000000E+00", "8.5000000E+00"]
    assert len(lines[2]) == 9 * 15


def test_green_kubo_integral_of_an_exponential(md):
    """kappa_from_acf on C(t) = C0 exp(-t / tc): the running integral tends to C0 tc, and the
    unit chain (10 J/mol per particle, Angstrom/ps, Angstrom^3, K) gives the SI value
    computed here independently."""
    import numpy as np
    c0, tc, dt = 2.5e4, 0.05, 1.0e-4           # internal units^2, ps, ps
    t = dt * np.arange(20000)
    k = md.kappa_from_acf(c0 * np.exp(-t / tc), dt, 15000.0, 300.0)
    assert k[0] == 0.0 and np.all(np.diff(k) >= 0) and k[1] > 0
    j_si = (10.0 / 6.02214076e23) * 100.0       # J m/s per internal unit of J
    expect = c0 * j_si ** 2 * tc * 1e-12 / (3 * 15000.0e-30 * 1.380649e-23 * 300.0 ** 2)
    assert abs(k[-1] - expect) <= 1e-6 * expect
