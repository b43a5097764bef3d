"""Phase structure function of the first screen of every realization
(simulations/structure_function.py:11-48 of pyAtmosphere), computed on the
device by pa_structure_function (the reference's O(N^3) numpy loop,
theory/phase_screens/sf.py:4-8), next to the model's closed form and its
numerical integral."""
import numpy as np

from .. import engine
from .measure import Measure
from .result import Result


def calculate_structure_function(channel, output) -> np.ndarray:
    """Row structure function of one screen averaged over its rows, lags
    1 .. N - 1 (simulations/structure_function.py:11-12)."""
    return engine.structure_function(output)


class StructureFunctionResult(Result):
    def __init__(self, channel, *args, **kwargs):
        measures = [Measure(channel, "phase_screen", calculate_structure_function)]
        super().__init__(*args, channel=channel, measures=measures, **kwargs)
        self.init_theoretical()

    def init_theoretical(self):
        channel = self.measures[0].channel
        channel.path.init_phase_screens()   # the screens' wavelength defaults through their channel
        ps = channel.path.phase_screen
        k = 2 * np.pi / ps.wvl   # the screen's own wavelength (it defaults to the source's)
        self.get_theoretical = ps.model.sf_phi(self.r, k, ps.thickness)
        self.get_numerical_theoretical = np.asarray(ps.model.sf_phi_numeric(self.r, k, ps.thickness))

    @property
    def structure_function(self) -> np.ndarray:
        return np.asarray(self.measures[0], dtype=np.float64).mean(axis=0)

    @property
    def r(self) -> np.ndarray:
        return np.arange(1, self.channel.grid.resolution[0]) * self.channel.grid.delta

    def plot_output(self):
        """Simulated row structure function against the closed form and the
        quadrature, on a log scale."""
        from . import plotting
        return plotting.curves([(self.r, self.structure_function, f"simulated (n = {len(self.measures[0])})"),
                                (self.r, self.get_theoretical, "MVK closed form"),
                                (self.r, self.get_numerical_theoretical, "MVK quadrature")],
                               "lag r, m", "$D_\\phi(r)$, rad$^2$", title="StructureFunctionResult", logy=True)
