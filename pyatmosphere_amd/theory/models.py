"""Refractive-index spectra (theory/models.py:10-91 of pyAtmosphere)."""
from dataclasses import dataclass

import numpy as np

from .atmosphere import get_r0


@dataclass
class Model:
    Cn2: float
    l0: float
    L0: float

    def psd_n_f(self, f):
        return self.psd_n(2 * np.pi * f)

    def psd_phi(self, kappa, k, thickness):
        """Phase spectrum of a slab: 2 pi k^2 dz Phi_n(kappa)."""
        return 2 * np.pi * k ** 2 * thickness * self.psd_n(kappa)

    def psd_phi_f(self, f, k, thickness):
        return 2 * np.pi * k ** 2 * thickness * self.psd_n_f(f)

    def sf_phi_numeric(self, r, k, thickness):
        """Phase structure function by quadrature of the spectrum
        (theory/models.py:25-32): 8 pi^2 * 2 pi k^2 dz * int Phi_n(2 pi f)(1 - J0(2 pi f r)) 2 pi f df."""
        import scipy.integrate
        import scipy.special
        coeff = 2 * np.pi * k ** 2 * thickness

        def integrand(f, ri):
            return self.psd_n(2 * np.pi * f) * (1 - scipy.special.j0(2 * np.pi * f * ri)) * 2 * np.pi * f
        return np.array([coeff * (2 * np.pi) * 2 * scipy.integrate.quad(integrand, 0, np.inf, args=(ri,),
                                                                        epsrel=1e-3)[0] * (2 * np.pi)
                         for ri in np.atleast_1d(r)])


class MVKModel(Model):
    def psd_n(self, kappa):
        """Modified von Karman: 0.033 Cn2 exp(-(kappa/km)^2) / (kappa^2 + k0^2)^(11/6),
        km = 5.92/l0, k0 = 2 pi/L0."""
        k0 = 2 * np.pi / self.L0
        km = 5.92 / self.l0
        return 0.033 * self.Cn2 * np.exp(-(kappa / km) ** 2) / (kappa ** 2 + k0 ** 2) ** (11 / 6)

    def sf_phi(self, r, k, thickness):
        """Closed-form MVK phase structure function."""
        k0 = 2 * np.pi / self.L0
        r0 = get_r0(self.Cn2, k, thickness)
        return 7.75 * r0 ** (-5 / 3) * self.l0 ** (-1 / 3) * r ** 2 * (
            1 / (1 + 2.03 * r ** 2 / self.l0 ** 2) ** (1 / 6) - 0.72 * (k0 * self.l0) ** (1 / 3))
