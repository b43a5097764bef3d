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


class _WindScreen:
    """Draws a new 'spectrum' at cache_clear and reuses it for every shift."""

    def __init__(self, rng):
        self.rng, self.base = rng, None

    def cache_clear(self):
        self.base = float(self.rng.integers(0, 1000))


class FakeWindChannel:
    def __init__(self):
        self.screen = _WindScreen(np.random.default_rng(4))
        self.path = type("WP", (), {"phase_screens": [self.screen], "positions": [1.0]})()
        self.pupil = None
        self.output = None

    def generator(self, pupil=False, shift=(0, 0), store_output=True, wind=False, **kw):
        assert wind
        yield np.float64(0), np.float64(0)
        self.output = np.float64(self.screen.base + shift[1])


def test_wind_time_series_reuses_the_realization(tmp_path):
    """A measure with `time` records one value per wind shift; every shift of
    one realization sees the same screens (cache_clear once per realization,
    simulations/simulation.py), and the CSV round trip keeps the lists."""
    from pyatmosphere_amd.simulations import WindResult
    ch = FakeWindChannel()
    times = (0.0, 0.5, 2.0)
    path = tmp_path / "wind.csv"
    r = WindResult(ch, [Measure(ch, "atmosphere", _value, name="v", time=times)], max_size=4, save_path=str(path))
    Simulation(results_list=[r]).run()
    data = r.measures[0].data
    assert len(data) == 4
    for row in data:
        assert row == [row[0] + t for t in times]
    assert len({row[0] for row in data}) > 1
    r.save_output()
    r2 = WindResult(ch, [Measure(ch, "atmosphere", _value, name="v", time=times)], save_path=str(path))
    assert r2.measures[0].data == data


def test_time_coherence_and_wander_correlations():
    """tc = Pearson correlation of eta at shift 0 with eta at each shift
    (scipy.stats.pearsonr); xx, yy, xy = 2 sqrt(<x0 xt>), 2 sqrt(<y0 yt>),
    2 sqrt(|<x0 yt>|) (simulations/wind.py:20-66)."""
    from scipy.stats import pearsonr
    from pyatmosphere_amd.simulations import TimeBWcorrSimulation, TimeCoherenceResult
    rng = np.random.default_rng(9)
    times = (0.0, 0.1, 1.0)
    base = rng.normal(size=200)
    eta = np.stack([base, base + 0.3 * rng.normal(size=200), rng.normal(size=200)], axis=1)
    tc = TimeCoherenceResult(None, times)
    tc.measures[0].data = eta.tolist()
    ref = [pearsonr(eta[:, 0], eta[:, i])[0] for i in range(3)]
    assert np.allclose(tc.tc, ref, rtol=1e-12, atol=1e-15) and tc.tc[0] == pytest.approx(1.0, abs=1e-15)
    x, y = rng.normal(size=(200, 3)) * 0.01, rng.normal(size=(200, 3)) * 0.01
    x[:, 1] = x[:, 0]
    bw = TimeBWcorrSimulation(None, times)
    bw.measures[0].data, bw.measures[1].data = x.tolist(), y.tolist()
    assert np.allclose(bw.xx[:2], 2 * np.sqrt((x[:, 0] ** 2).mean()), rtol=1e-12)
    with np.errstate(invalid="ignore"):   # negative <y0 yt> give NaN, as in the reference
        assert bw.yy[0] == pytest.approx(2 * np.sqrt((y[:, 0] ** 2).mean()), rel=1e-12)
    assert np.allclose(bw.xy, 2 * np.sqrt(np.abs((x[:, :1] * y).mean(axis=0))), rtol=1e-12)


def test_si_result_formula():
    """SI = <I^2> / <I>^2 - 1 per position over the realizations, positions =
    screen positions then the receiver (simulations/si.py:26-36)."""
    from pyatmosphere_amd.simulations import SIResult
    ch = FakeChannel()
    ch.path.length = 3.0
    r = SIResult(ch, theoretical_functions=())
    rng = np.random.default_rng(2)
    data = rng.exponential(size=(500, 3))
    r.measures[0].data = data.tolist()
    assert np.allclose(r.positions, [1.0, 2.0, 3.0])
    assert np.allclose(r.si, data.var(axis=0) / data.mean(axis=0) ** 2, rtol=1e-12)
