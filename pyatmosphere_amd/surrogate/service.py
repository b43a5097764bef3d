"""Transmittance PDTs on demand: the surrogate where it is confident, the full
simulation where it is not.

`transmittance_pdts` answers each (channel parameters, aperture radius)
request with the surrogate ensemble's 64-bin PDT when the inputs are inside
the training domain and the members agree (SurrogateEnsemble.pdt), and
otherwise runs the reference's own Monte Carlo for that channel,
Simulation([PDTResult]) (simulations/pdt.py:11-26; simulations/simulation.py:
89-138), on a Channel built from the parameters with the grid rule of the
training data, and bins its transmittance samples the same way.
"""
from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import data as sdata
from .model import NB, SurrogateEnsemble, encode


def build_channel(p: Dict, radius: float):
    """A pyAtmosphere-style Channel for parameters p (log10_cn2, L, w0, wvl) and
    one circular aperture: the training data's grid rule (channel_grid), MVK
    screens with l0 = 6 mm, L0 = 1 km and 1024 log-polar points."""
    from .. import (Channel, CirclePupil, GaussianSource, IdenticalPhaseScreensPath, MVKModel, RandLogPolarGrid,
                    RectGrid, SSPhaseScreen)
    n, delta, S = sdata.channel_grid(p)
    screen = SSPhaseScreen(model=MVKModel(Cn2=10 ** p["log10_cn2"], l0=sdata.l0, L0=sdata.L0),
                           f_grid=RandLogPolarGrid(points=sdata.POINTS, f_min=1 / (15 * sdata.L0), f_max=2 / sdata.l0))
    return Channel(grid=RectGrid(resolution=n, delta=delta),
                   source=GaussianSource(wvl=p["wvl"], w0=p["w0"], F0=np.inf),
                   path=IdenticalPhaseScreensPath(phase_screen=screen, length=p["L"], count=S),
                   pupil=CirclePupil(radius=radius))


def simulated_pdt(p: Dict, radius: float, realizations: int, seed: int) -> np.ndarray:
    """Bin probabilities (NB,) of Simulation([PDTResult]) for one channel and aperture."""
    from ..simulations import PDTResult, Simulation
    ch = build_channel(p, radius)
    res = PDTResult(ch, max_size=realizations)
    Simulation([res], rng="device", seed=seed).run()
    eta = np.asarray(res.measures[0].data, dtype=np.float64)
    h, _ = np.histogram(eta, bins=NB, range=(0.0, 1.0))
    return h / max(len(eta), 1)


def transmittance_pdts(requests: Sequence[Tuple[Dict, float]], ensemble: SurrogateEnsemble,
                       realizations: int = 2000, seed: int = 0) -> Tuple[np.ndarray, List[str]]:
    """PDTs (M, NB) for M (channel parameters, aperture radius) requests, and
    for each the source that answered it: "surrogate" or "simulation"."""
    x = np.stack([encode(p["log10_cn2"], p["L"], p["w0"], a, p["wvl"]) for p, a in requests])
    pdt, _, _, fallback = ensemble.pdt(x)
    out = np.array(pdt, dtype=np.float64)
    source = ["surrogate"] * len(requests)
    for i in np.nonzero(fallback)[0]:
        p, a = requests[i]
        out[i] = simulated_pdt(p, a, realizations, seed + int(i))
        source[i] = "simulation"
    return out, source
