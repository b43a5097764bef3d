"""Probability distribution of transmittance (simulations/pdt.py:11-66 of pyAtmosphere)."""
from functools import partial

import numpy as np

from ..measures import eta, mean_x, mean_y
from .measure import Measure
from .result import Result

PDT_BINS = 64


def pdt_histogram(samples, bins=PDT_BINS):
    """Density histogram of transmittance samples over [0, 1] (64 bins)."""
    h, edges = np.histogram(np.asarray(samples, dtype=np.float64), bins=bins, range=(0.0, 1.0), density=True)
    return h, edges


class PDTResult(Result):
    def __init__(self, channel, pupils: list = None, **kwargs):
        self.pupil_shift = (0, 0)
        self.pupils = pupils or [channel.pupil]
        measures = kwargs.pop("measures", [
            Measure(channel, "atmosphere", partial(self.append_pupil, pupil), eta, name=f"{pupil.radius}")
            for pupil in self.pupils])
        super().__init__(channel, measures, **kwargs)

    def append_pupil(self, pupil, channel, output):
        init_pupil = channel.pupil
        channel.pupil = pupil
        output = channel.pupil.output(output, shift=self.pupil_shift)
        channel.pupil = init_pupil
        return output

    def histogram(self, index=0, bins=PDT_BINS):
        return pdt_histogram(self.measures[index].data, bins)

    def plot_output(self):
        """The 64-bin transmittance density of every pupil on one axes."""
        from . import plotting
        return plotting.densities([m.data for m in self.measures[-len(self.pupils):]],
                                  [f"a = {p.radius:.3g} m" for p in self.pupils], title=type(self).__name__)


class TrackedPDTResult(PDTResult):
    """Pupil re-centred on each realization's beam centroid (pdt.py:51-66)."""

    def __init__(self, channel, pupils: list = None, **kwargs):
        pupils = pupils or [channel.pupil]
        beam_measures = [Measure(channel, "atmosphere", mean_x), Measure(channel, "atmosphere", mean_y)]
        pdt_measures = [
            Measure(channel, "atmosphere", self.set_pupil_position, partial(self.append_pupil, pupil), eta,
                    name=f"{pupil.radius}")
            for pupil in pupils]
        super().__init__(channel, pupils=pupils, measures=beam_measures + pdt_measures, **kwargs)

    def set_pupil_position(self, channel, output):
        self.pupil_shift = (self.measures[0].iteration_data, self.measures[1].iteration_data)
        return output
