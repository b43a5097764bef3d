"""Frozen-flow (wind) time-series results (simulations/wind.py:13-66 of
pyAtmosphere).

A measure built with `time=[t0, t1, ...]` records one value per time for each
realization: the Simulation loop draws the realization's screens once and
re-evaluates them with the coordinates shifted by (0, t) (the wind path,
phase_screens.py:108-136 with `wind=True`; on the device the shift is folded
into the spectral amplitudes), so the columns of a measure's data are the same
atmosphere seen after the wind has moved it by t metres.
"""
from typing import List

import numpy as np

from ..measures import eta, mean_x, mean_y
from .measure import Measure
from .result import Result


def _columns(measure) -> np.ndarray:
    """(realizations, times) array of a time-series measure."""
    return np.asarray(measure, dtype=np.float64)


def _pearson(a: np.ndarray, b: np.ndarray) -> float:
    """Pearson correlation coefficient of two samples (scipy.stats.pearsonr's
    statistic)."""
    a = a - a.mean()
    b = b - b.mean()
    return float((a * b).sum() / np.sqrt((a * a).sum() * (b * b).sum()))


class WindResult(Result):
    """Result whose measures hold one list per realization (one value per
    time); the CSV keeps each list as its text "[v0, v1, ...]"
    (simulations/wind.py:13-17)."""

    def load_output(self):
        import pandas as pd
        for measures, column in zip(self.measures, pd.read_csv(self.save_path).T.values):
            measures.data = [[float(v) for v in cell.strip()[1:-1].split(",")]
                             for cell in column if isinstance(cell, str)]


class TimeCoherenceResult(WindResult):
    """Transmittance eta through the channel's pupil at each wind shift; `tc`
    is its correlation between shift 0 and each shift (simulations/wind.py:20-35)."""

    def __init__(self, channel, time, *args, **kwargs):
        measures = [Measure(channel, "pupil", eta, time=time)]
        super().__init__(*args, channel=channel, measures=measures, **kwargs)

    @property
    def tc(self) -> List[float]:
        d = _columns(self.measures[0])
        return [_pearson(d[:, 0], d[:, i]) for i in range(d.shape[1])]

    def plot_output(self):
        """Correlation of eta with its value at the first shift, per shift."""
        from . import plotting
        if len(self.measures[0]) < 3:
            return None
        return plotting.curves([(self.measures[0].time, self.tc, f"n = {len(self.measures[0])}")],
                               "wind shift, m", "corr($\\eta_0$, $\\eta_t$)", title="TimeCoherenceResult",
                               ylim=(0.0, 1.0))


class TimeBWcorrSimulation(WindResult):
    """Beam-centroid (mean_x, mean_y) at each wind shift; `xx`, `yy`, `xy` are
    2 sqrt(<x_0 x_t>), 2 sqrt(<y_0 y_t>), 2 sqrt(|<x_0 y_t>|) over the
    realizations (simulations/wind.py:38-66)."""

    def __init__(self, channel, time, *args, **kwargs):
        measures = [Measure(channel, "atmosphere", mean_x, time=time),
                    Measure(channel, "atmosphere", mean_y, time=time)]
        super().__init__(*args, channel=channel, measures=measures, **kwargs)

    @property
    def xx(self) -> np.ndarray:
        x = _columns(self.measures[0])
        return 2 * np.sqrt((x[:, :1] * x).mean(axis=0))

    @property
    def yy(self) -> np.ndarray:
        y = _columns(self.measures[1])
        return 2 * np.sqrt((y[:, :1] * y).mean(axis=0))

    @property
    def xy(self) -> np.ndarray:
        x, y = _columns(self.measures[0]), _columns(self.measures[1])
        return 2 * np.sqrt(np.abs((x[:, :1] * y).mean(axis=0)))

    def plot_output(self):
        """The three centroid correlation widths against the wind shift."""
        from . import plotting
        t = self.measures[0].time
        return plotting.curves([(t, self.xx, "xx"), (t, self.yy, "yy"), (t, self.xy, "|xy|")],
                               "wind shift, m", "2 sqrt(<c_0 c_t>), m", title="TimeBWcorrSimulation",
                               scatter=True)
