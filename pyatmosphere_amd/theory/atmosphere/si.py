"""On-axis scintillation index of a Gaussian beam in moderate-to-strong
turbulence with inner and outer scale: Andrews & Phillips, Laser Beam
Propagation through Random Media, 2nd ed., p. 356 eq. (114), as pyAtmosphere's
get_SI_andrews_strong (theory/atmosphere/si.py:94-146)."""
import numpy as np

from . import get_rytov2


def get_SI_andrews_strong(length, model, gaussian_beam):
    th, lam = gaussian_beam.get_theta(length), gaussian_beam.get_Lambda(length)
    thb = 1 - th
    k = gaussian_beam.k
    Ql = 10.89 * length / k / model.l0 ** 2
    Q0 = 64 * np.pi ** 2 * length / k / model.L0 ** 2
    sr2 = get_rytov2(model.Cn2, k, length)
    a, b = 1 + 2 * th, 3 + 2 * lam * Ql
    phi1, phi2 = np.arctan(2 * lam / a), np.arctan(a * Ql / b)
    den = a ** 2 * Ql ** 2 + b ** 2
    # small-scale (Rytov) on-axis variance with inner scale
    sg2 = 3.86 * sr2 * (
        0.4 * (a ** 2 + (2 * lam + 3 / Ql) ** 2) ** (11 / 12) / np.sqrt(a ** 2 + 4 * lam ** 2)
        * (np.sin(11 / 6 * phi2 + phi1) + 2.61 * den ** -0.25 * np.sin(4 / 3 * phi2 + phi1)
           - 0.52 * den ** (-7 / 24) * np.sin(5 / 4 * phi2 + phi1))
        - 13.40 * lam * Ql ** (-11 / 6) / (a ** 2 + 4 * lam ** 2)
        - 11 / 6 * ((1 + 0.31 * lam * Ql) ** (5 / 6) + 1.1 * (1 + 0.27 * lam * Ql) ** (1 / 3)
                    - 0.19 * (1 + 0.24 * lam * Ql) ** 0.25) * Ql ** (-5 / 6))
    # large-scale log-amplitude variance: inner-scale term minus its outer-scale cutoff
    poly = 1 / 3 - thb / 2 + thb ** 2 / 5
    eta_x = 1 / (0.38 / (1 - 3.21 * thb + 5.29 * thb ** 2)
                 + 0.47 * sr2 * Ql ** (1 / 6) * (poly / (1 + 2.2 * thb)) ** (6 / 7))

    def large_scale(eta):
        r = eta / (eta + Ql)
        return 0.49 * sr2 * poly * (eta * Ql / (eta + Ql)) ** (7 / 6) * (1 + 1.75 * np.sqrt(r) - 0.25 * r ** (7 / 12))

    ln_x = large_scale(eta_x) - large_scale(eta_x * Q0 / (eta_x + Q0))
    ln_y = 0.51 * sg2 / (1 + 0.69 * sg2 ** (6 / 5)) ** (5 / 6)
    return np.exp(ln_x + ln_y) - 1
