"""MI355X-native split-step phase-screen path with pyAtmosphere's interface.

    import pyatmosphere_amd as pa
    ch = pa.QuickChannel(Cn2=1e-15, length=10e3, aperture_radius=0.12)
    u = ch.run()                      # device field (torch complex64)
    pa.measures.eta(ch, output=u)

The compute runs in libpa_hip.so (hand-written HIP kernels for gfx950) behind
the reference's constructors; see include/pa_hip.h for the C ABI.
"""
from . import gpu, measures, simulations
from .channels import Channel, QuickChannel
from .grids import Grid, RandLogPolarGrid, RectGrid
from .pathes import AbstractPath, IdenticalPhaseScreensPath, PhaseScreensPath, VacuumPath
from .phase_screens import FFTPhaseScreen, PhaseScreen, SSPhaseScreen, SUPhaseScreen
from .pupils import CirclePupil
from .sources import GaussianSource, PlaneSource, Source
from .theory.models import Model, MVKModel

__all__ = ["Channel", "QuickChannel"]
