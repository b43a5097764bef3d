"""Independent pins of the package's theory closed forms (the statistical
judges): each is checked against a limit or a quadrature it must reduce to,
chosen so that a dropped term, a wrong exponent or a wrong sign fails.

* get_rytov2 (theory/atmosphere/__init__.py:5-6): the main.ipynb value 27.7.
* get_SI_andrews_strong (theory/atmosphere/si.py:94-146): in weak turbulence
  (sigma_R^2 Q_l^(1/6) << 1) with a vanishing inner scale and an infinite outer
  scale, Andrews & Phillips eq. 114 reduces to the Rytov variance of a plane
  wave (w0 -> infinity: Theta = 1, Lambda = 0), 1.23 Cn2 k^(7/6) L^(11/6), and of
  a spherical wave (w0 -> 0: Theta = Lambda = 0), 0.5 Cn2 k^(7/6) L^(11/6).
* MVKModel.sf_phi (theory/models.py:87-91, with get_r0 of
  theory/atmosphere/__init__.py:9-11): against the quadrature of the same
  spectrum, Model.sf_phi_numeric (theory/models.py:25-32).
* Model.sf_phi_numeric: in the Kolmogorov limit, 6.88 (r / r0)^(5/3).
"""
import numpy as np
import pytest

from pyatmosphere_amd.theory.atmosphere import get_r0, get_rytov2
from pyatmosphere_amd.theory.atmosphere.si import get_SI_andrews_strong
from pyatmosphere_amd.theory.models import MVKModel
from pyatmosphere_amd.theory.sources import GaussianBeam


def test_rytov_main_notebook():
    assert round(get_rytov2(5e-16, 2 * np.pi / 808e-9, 50e3), 1) == 27.7


WEAK = [(1e-20, 1e3, 808e-9), (1e-21, 5e3, 1550e-9), (1e-23, 2e4, 808e-9)]


@pytest.mark.parametrize("Cn2,L,wvl", WEAK)
def test_si_andrews_plane_wave_weak_limit(Cn2, L, wvl):
    """w0 = 1 km (Lambda ~ 1e-9), l0 = 1e-7 m, L0 = 1e9 m, sigma_R^2 Q_l^(1/6) < 1e-4:
    the formula's SI equals sigma_R^2 within 2% (it gives 1.014 sigma_R^2)."""
    model = MVKModel(Cn2=Cn2, l0=1e-7, L0=1e9)
    beam = GaussianBeam(wvl=wvl, w0=1e3, F0=np.inf)
    k = 2 * np.pi / wvl
    sr2 = 1.23 * Cn2 * k ** (7 / 6) * L ** (11 / 6)
    ql = 10.89 * L / k / model.l0 ** 2
    assert sr2 * ql ** (1 / 6) < 1e-4
    si = get_SI_andrews_strong(L, model, beam)
    assert abs(si / sr2 - 1) < 0.02, si / sr2


@pytest.mark.parametrize("Cn2,L,wvl", WEAK)
def test_si_andrews_spherical_wave_weak_limit(Cn2, L, wvl):
    """w0 = 10 um (Theta, Lambda ~ 1e-7 or below): the spherical-wave Rytov
    variance 0.5 Cn2 k^(7/6) L^(11/6), within 5% (the fit gives 0.963 of it)."""
    model = MVKModel(Cn2=Cn2, l0=1e-7, L0=1e9)
    beam = GaussianBeam(wvl=wvl, w0=1e-5, F0=np.inf)
    k = 2 * np.pi / wvl
    beta2 = 0.5 * Cn2 * k ** (7 / 6) * L ** (11 / 6)
    si = get_SI_andrews_strong(L, model, beam)
    assert abs(si / beta2 - 1) < 0.05, si / beta2


def test_si_andrews_outer_scale_lowers_the_index():
    """A finite outer scale removes large-scale (refractive) power: the index
    falls as L0 shrinks, and moves towards the no-outer-scale value as L0 grows."""
    beam = GaussianBeam(wvl=808e-9, w0=0.09, F0=np.inf)
    vals = [get_SI_andrews_strong(1e4, MVKModel(Cn2=1e-15, l0=3e-3, L0=L0), beam) for L0 in (1.0, 10.0, 1e2, 1e3, 1e6)]
    assert all(a < b for a, b in zip(vals, vals[1:])), vals


def test_mvk_sf_phi_against_quadrature():
    """Closed form vs quadrature of the MVK spectrum for lags 1 mm .. 1 m
    (they agree to 0.98-1.02)."""
    model = MVKModel(Cn2=1e-15, l0=6e-3, L0=1e3)
    k = 2 * np.pi / 808e-9
    r = np.logspace(-3, 0, 13)
    ratio = model.sf_phi(r, k, 1e4) / model.sf_phi_numeric(r, k, 1e4)
    assert np.all(np.abs(ratio - 1) < 0.03), ratio


def test_sf_phi_numeric_kolmogorov_limit():
    """l0 -> 0, L0 -> infinity: D(r) = 6.88 (r / r0)^(5/3) with the plane-wave r0
    (0.423 k^2 Cn2 L)^(-3/5), within 5% for lags 3 mm .. 10 cm."""
    model = MVKModel(Cn2=1e-15, l0=1e-9, L0=1e9)
    k = 2 * np.pi / 808e-9
    r = np.logspace(-2.5, -1, 4)
    d = model.sf_phi_numeric(r, k, 1e4)
    kol = 6.88 * (r / get_r0(1e-15, k, 1e4)) ** (5 / 3)
    assert np.all(np.abs(d / kol - 1) < 0.05), d / kol
