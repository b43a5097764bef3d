"""Channel descriptions of the standard workloads (synthetic parameters; the
spectra are drawn at run time): the 2048^2 / 50 km / 5-screen channel, the
512^2 / 10 km / 5-screen channel, and the mixed-channel Monte Carlo workload
of 256 slots over three grid sizes."""
from typing import List

import numpy as np

from . import engine


def _mvk_psd(base, wvl, thickness, Cn2, l0, L0):
    return engine.annulus_psd(base, 2 * np.pi / wvl, thickness,
                              lambda k: engine.mvk_psd_n(k, Cn2, l0, L0)).astype(np.float32)


def _middle_positions(length, count):
    """Screen positions in the middle of each slab (pathes.py:86-87)."""
    return (np.arange(count) + 0.5) * length / count


def channel_2048_50km(pupil_radius: float = 0.2) -> engine.ChannelSpec:
    """2048^2 grid, delta 1.5 mm, collimated Gaussian beam (808 nm, w0 0.12 m),
    50 km with 5 SS/MVK screens (Cn2 5e-16, l0 6 mm, L0 1 km), 1024 log-polar
    points over [1/15000, 2/6e-3] 1/m, circular pupil of radius 0.2 m."""
    base = engine.log_polar_base(1024, 1 / 15000, 2 / 6e-3)
    L, S, wvl = 50e3, 5, 808e-9
    return engine.ChannelSpec(n=2048, delta=1.5e-3, wvl=wvl, w0=0.12, F0=np.inf, length=L,
                              positions=_middle_positions(L, S), points=1024, base=base,
                              psd=_mvk_psd(base, wvl, L / S, 5e-16, 6e-3, 1e3), pupils=[(pupil_radius, 0.0, 0.0)])


def channel_512_10km() -> engine.ChannelSpec:
    """512^2 grid, delta 1.5 mm, collimated Gaussian beam (808 nm, w0 0.09 m),
    10 km with 5 SS/MVK screens (Cn2 1e-15, l0 6 mm, L0 1 km), 1024 log-polar
    points over [1/15000, 2/6e-3] 1/m, circular pupil of radius 0.12 m."""
    base = engine.log_polar_base(1024, 1 / 15000, 2 / 6e-3)
    L, S, wvl = 10e3, 5, 808e-9
    return engine.ChannelSpec(n=512, delta=1.5e-3, wvl=wvl, w0=0.09, F0=np.inf, length=L,
                              positions=_middle_positions(L, S), points=1024, base=base,
                              psd=_mvk_psd(base, wvl, L / S, 1e-15, 6e-3, 1e3), pupils=[(0.12, 0.0, 0.0)])


def mixed_channels(slots: int = 256, seed: int = 2024, pupil_radius: float = 0.1) -> List[engine.ChannelSpec]:
    """One channel per slot, drawn from numpy default_rng(seed): grid n in
    {512, 1024, 2048}, delta in {1, 1.5, 2} mm, 3..10 screens, log10 Cn2 in
    [-16, -14], L in [1, 50] km, lambda in {808, 1550} nm, w0 in [0.02, 0.2] m
    (MVK l0 6 mm, L0 1 km, 1024 log-polar points over [1/15000, 2/6e-3] 1/m,
    one circular pupil)."""
    rng = np.random.default_rng(seed)
    base = engine.log_polar_base(1024, 1 / 15000, 2 / 6e-3)
    specs = []
    for _ in range(slots):
        n = int(rng.choice([512, 1024, 2048]))
        delta = float(rng.choice([1e-3, 1.5e-3, 2e-3]))
        S = int(rng.integers(3, 11))
        Cn2 = 10 ** rng.uniform(-16, -14)
        L = rng.uniform(1e3, 50e3)
        wvl = float(rng.choice([808e-9, 1550e-9]))
        w0 = rng.uniform(0.02, 0.2)
        specs.append(engine.ChannelSpec(n=n, delta=delta, wvl=wvl, w0=w0, F0=np.inf, length=L,
                                        positions=_middle_positions(L, S), points=1024, base=base,
                                        psd=_mvk_psd(base, wvl, L / S, Cn2, 6e-3, 1e3),
                                        pupils=[(pupil_radius, 0.0, 0.0)]))
    return specs
