"""Monte Carlo simulations over channels (simulations/__init__.py of pyAtmosphere)."""
from .beam import BeamResult
from .measure import Measure
from .pdt import PDTResult, TrackedPDTResult, pdt_histogram
from .result import Result
from .si import SIResult, intensity_at_center
from .simulation import Simulation
from .structure_function import StructureFunctionResult, calculate_structure_function
from .wind import TimeBWcorrSimulation, TimeCoherenceResult, WindResult

__all__ = ["Result", "Simulation", "Measure", "BeamResult", "PDTResult", "TrackedPDTResult", "pdt_histogram",
           "WindResult", "TimeCoherenceResult", "TimeBWcorrSimulation", "SIResult", "intensity_at_center",
           "StructureFunctionResult", "calculate_structure_function"]
