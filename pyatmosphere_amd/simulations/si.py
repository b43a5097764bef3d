"""On-axis scintillation index after every screen (simulations/si.py:11-36 of
pyAtmosphere): the intensity at the grid origin is recorded after each screen
and at the receiver (a "propagation" measure), and SI = <I^2> / <I>^2 - 1 per
position, next to the theory's on-axis SI."""
import numpy as np

from ..theory.atmosphere.si import get_SI_andrews_strong
from .measure import Measure
from .result import Result


def intensity_at_center(channel, output) -> float:
    """|u|^2 at the grid origin (simulations/si.py:11-12); one element of the
    device field is read back."""
    i, j = channel.grid.origin_index
    v = complex(output[i, j].item()) if hasattr(output[i, j], "item") else complex(output[i, j])
    return float(abs(v) ** 2)


class SIResult(Result):
    def __init__(self, channel, theoretical_functions=(get_SI_andrews_strong,), *args, **kwargs):
        measures = [Measure(channel, "propagation", intensity_at_center)]
        super().__init__(*args, channel=channel, measures=measures, **kwargs)
        self.set_theoretical_functions(*theoretical_functions)

    def set_theoretical_functions(self, *theoretical_functions):
        self.theoretical_functions = theoretical_functions
        self.theoretical_si = [f(self.positions, self.channel.path.phase_screen.model, self.channel.source)
                               for f in theoretical_functions]

    @property
    def intensities_at_center(self) -> np.ndarray:
        """(realizations, screens + 1): after each screen, then at the receiver."""
        return np.asarray(self.measures[0], dtype=np.float64)

    @property
    def positions(self) -> np.ndarray:
        return np.array(list(self.channel.path.positions) + [self.channel.path.length])

    @property
    def si(self) -> np.ndarray:
        i = self.intensities_at_center
        return (i ** 2).mean(axis=0) / i.mean(axis=0) ** 2 - 1

    def plot_output(self):
        """Simulated on-axis SI at each screen and at the receiver, beside each
        theory curve."""
        from . import plotting
        series = [(self.positions, self.si, f"simulated (n = {len(self.measures[0])})")]
        series += [(self.positions, t, f.__name__) for f, t in zip(self.theoretical_functions, self.theoretical_si)]
        return plotting.curves(series, "z, m", "on-axis scintillation index", title="SIResult")
