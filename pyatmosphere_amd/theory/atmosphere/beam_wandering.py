"""Beam-wander radius of a Gaussian beam, Klyatskin & Kon, Radiophys. Quantum
Electron. 15, 1056 (1972), in the form of pyAtmosphere's get_r_bw
(theory/atmosphere/beam_wandering.py:5-15): a double integral over the path
coordinate z and the spatial frequency kappa."""
import numpy as np
import scipy.integrate


def get_r_bw(L, model, gaussian_beam):
    """Root mean square beam-wander radius at distance L (2-D: sigma_BW,x * sqrt 2)."""
    k = gaussian_beam.k
    kappa0 = 2 * np.pi / model.L0

    def phase_sf(rho):
        # spherical-wave phase structure function of the MVK spectrum over the whole path
        return 1.09 * model.Cn2 * k ** 2 * L * model.l0 ** (-1 / 3) * rho ** 2 * (
            (1 + (rho / model.l0) ** 2) ** (-1 / 6) - 0.72 * (kappa0 * model.l0) ** (1 / 3))

    def integrand(z, kappa):
        w = gaussian_beam.get_w(z)
        return (4 * np.pi ** 2 * (L - z) ** 2 * kappa ** 3 * model.psd_n(kappa)
                * np.exp(-(kappa * w) ** 2 / 4 - np.pi * phase_sf(kappa * z / k)))

    return float(np.sqrt(scipy.integrate.dblquad(integrand, 0, np.inf, 0, L)[0]))
