"""Beam wander / long-term / short-term width statistics
(simulations/beam.py:14-80 of pyAtmosphere)."""
from typing import Sequence, Tuple

import numpy as np

from ..measures import I, mean_x, mean_x2, mean_xy, mean_y, mean_y2
from .. import engine
from .measure import Measure
from .result import Result


def _sqrt_mean_with_error(v):
    """sqrt(mean v) and its error std(ddof=1)/sqrt(n)/2/sqrt(mean) (beam.py:49-71)."""
    v = np.asarray(v, dtype=np.float64)
    m = np.sqrt(v.mean())
    return m, v.std(ddof=1) / np.sqrt(len(v)) / 2 / m


class BeamResult(Result):
    def __init__(self, channel, **kwargs):
        measures = [
            Measure(channel, "atmosphere", mean_x),
            Measure(channel, "atmosphere", mean_y),
            Measure(channel, "atmosphere", mean_x2),
            Measure(channel, "atmosphere", mean_xy),
            Measure(channel, "atmosphere", mean_y2),
            Measure(channel, "atmosphere", self.mean_x2_r, name="mean_x2_r"),
        ]
        super().__init__(channel, measures, **kwargs)

    def mean_x2_r(self, channel, output):
        """Second moment along the centroid direction of this realization
        (beam.py:26-33); equals cos^2 <x2> + 2 cos sin <xy> + sin^2 <y2>."""
        mx, my = self.measures[0].iteration_data, self.measures[1].iteration_data
        r0 = np.sqrt(mx ** 2 + my ** 2)
        c, s = mx / r0, my / r0
        m = engine.field_moments(output, channel.grid.delta)[0]
        return float(c * c * m[3] + 2 * c * s * m[4] + s * s * m[5])

    @property
    def bw2(self) -> Sequence[float]:
        return np.asarray(self.measures[0]) ** 2

    @property
    def lt2(self) -> Sequence[float]:
        return 4 * np.asarray(self.measures[2])

    @property
    def st2(self) -> Sequence[float]:
        return self.lt2 - 4 * self.bw2

    @property
    def bw(self) -> Tuple[float, float]:
        return _sqrt_mean_with_error(self.bw2)

    @property
    def lt(self) -> Tuple[float, float]:
        return _sqrt_mean_with_error(self.lt2)

    @property
    def st(self) -> Tuple[float, float]:
        return _sqrt_mean_with_error(self.st2)

    def print_output(self):
        bw, lt, st = self.bw, self.lt, self.st
        print(f"sigma_BW_x = {bw[0]:.1e} +- {bw[1]:.1e}")
        print(f"sigma_LT_x = {lt[0]:.1e} +- {lt[1]:.1e}")
        print(f"W_ST = {st[0]:.1e} +- {st[1]:.1e}")
        print(f"Count of measures: {len(self.measures[0])}")
