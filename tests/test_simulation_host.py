"""Monte Carlo driver bookkeeping on the host (no GPU): measure grouping, max_size,
KeyboardInterrupt handling and the CSV checkpoint/resume of simulations/result.py."""
import numpy as np
import pytest

from pyatmosphere_amd.simulations import BeamResult, Measure, PDTResult, Result, Simulation, pdt_histogram
from pyatmosphere_amd.simulations.simulation import _classify
from pyatmosphere_amd import measures as M


class _Screen:
    def cache_clear(self):
        pass


class _Path:
    def __init__(self):
        self.phase_screens = [_Screen(), _Screen()]
        self.positions = [1.0, 2.0]


class FakeChannel:
    """Yields numpy 'fields' whose value is a running counter (path generator shape)."""

    def __init__(self):
        self.path = _Path()
        self.pupil = None
        self.output = None
        self.calls = 0

    def generator(self, pupil=False, store_output=True, **kw):
        self.calls += 1
        for i in range(2):
            yield np.float64(self.calls * 10 + i), np.float64(-1)
        self.output = np.float64(self.calls)


def _value(channel, output):
    return float(output)


def test_generic_loop_max_size():
    ch = FakeChannel()
    m = Measure(ch, "atmosphere", _value, max_size=5)
    p = Measure(ch, "propagation", _value, max_size=3)
    Simulation(measures_list=[m, p]).run()
    assert m.data == [1.0, 2.0, 3.0, 4.0, 5.0]
    assert len(p.data) == 3 and p.data[0] == [10.0, 11.0, 1.0]


def test_keyboard_interrupt_saves_and_resumes(tmp_path):
    ch = FakeChannel()
    path = tmp_path / "res.csv"

    def stop(channel, output):
        if float(output) >= 4:
            raise KeyboardInterrupt
        return float(output) * 1.5

    r = Result(ch, [Measure(ch, "atmosphere", stop, name="v")], max_size=10, save_path=str(path))
    Simulation(results_list=[r]).run(save_step=1)
    assert r.measures[0].data == [1.5, 3.0, 4.5]
    r2 = Result(ch, [Measure(ch, "atmosphere", stop, name="v")], max_size=10, save_path=str(path))
    assert r2.measures[0].data == pytest.approx([1.5, 3.0, 4.5], rel=1e-3)  # '{:.3e}' round trip
    assert open(path).readline().strip() == "v"


def test_classify_standard_measures():
    ch = FakeChannel()
    ch.pupil = type("P", (), {"radius": 0.1})()
    beam = BeamResult(ch)
    kinds = [_classify(m) for m in beam.measures]
    assert kinds == [("moment", i) for i in (1, 2, 3, 4, 5, 6)]
    pdt = PDTResult(ch)
    assert _classify(pdt.measures[0]) == ("pupil", 0.1, 0.0, 0.0)
    assert _classify(Measure(ch, "atmosphere", _value)) is None
    assert _classify(Measure(ch, "atmosphere", M.eta)) == ("moment", 0)


def test_beam_summaries_match_reference_formulas():
    ch = FakeChannel()
    beam = BeamResult(ch)
    rng = np.random.default_rng(0)
    x = rng.normal(0, 0.04, 500)
    x2 = rng.uniform(0.01, 0.02, 500)
    beam.measures[0].data = list(x)
    beam.measures[2].data = list(x2)
    bw2 = x ** 2
    assert np.isclose(beam.bw[0], np.sqrt(bw2.mean()))
    assert np.isclose(beam.bw[1], bw2.std(ddof=1) / np.sqrt(500) / 2 / np.sqrt(bw2.mean()))
    st2 = 4 * x2 - 4 * bw2
    assert np.isclose(beam.st[0], np.sqrt(st2.mean()))


def test_pdt_histogram_density():
    h, edges = pdt_histogram(np.linspace(0.001, 0.999, 1000))
    assert len(h) == 64 and np.isclose((h * np.diff(edges)).sum(), 1.0)
