"""Turbulence closed forms used as statistical judges
(theory/atmosphere/__init__.py:5-11 of pyAtmosphere)."""


def get_rytov2(Cn2, k, length):
    """Rytov variance 1.23 Cn2 k^(7/6) L^(11/6)."""
    return 1.23 * Cn2 * k ** (7 / 6) * length ** (11 / 6)


def get_r0(Cn2, k, length):
    """Plane-wave coherence diameter (0.423 k^2 Cn2 L)^(-3/5)."""
    return (0.423 * k ** 2 * Cn2 * length) ** (-3 / 5)
