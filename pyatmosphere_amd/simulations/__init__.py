"""Monte Carlo simulations over channels (simulations/__init__.py of pyAtmosphere)."""
from .beam import BeamResult
from .measure import Measure
from .pdt import PDTResult, TrackedPDTResult, pdt_histogram
from .result import Result
from .simulation import Simulation
from .wind import TimeBWcorrSimulation, TimeCoherenceResult, WindResult

__all__ = ["Result", "Simulation", "Measure", "BeamResult", "PDTResult", "TrackedPDTResult", "pdt_histogram",
           "WindResult", "TimeCoherenceResult", "TimeBWcorrSimulation"]
