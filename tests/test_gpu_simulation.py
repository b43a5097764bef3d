"""Monte Carlo driver on the device: the fused batched path equals the generic
per-realization loop (same seeded draws) and the FP64 oracle."""
import numpy as np
import pytest

import pyatmosphere_amd as pa
from oracle import fp64 as o
from pyatmosphere_amd.simulations import BeamResult, PDTResult, Simulation, TrackedPDTResult

pytestmark = pytest.mark.gpu


def _channel(n=256, delta=0.002, L=5e3, S=3, P=128, Cn2=2e-15, w0=0.04, radius=0.05):
    return pa.Channel(
        grid=pa.RectGrid(resolution=n, delta=delta),
        source=pa.GaussianSource(wvl=808e-9, w0=w0, F0=np.inf),
        path=pa.IdenticalPhaseScreensPath(
            phase_screen=pa.SSPhaseScreen(model=pa.MVKModel(Cn2=Cn2, l0=6e-3, L0=1e3),
                                          f_grid=pa.RandLogPolarGrid(points=P, f_min=1 / 15000, f_max=2 / 6e-3)),
            length=L, count=S),
        pupil=pa.CirclePupil(radius=radius))


def _results(ch, n):
    return [BeamResult(ch, max_size=n),
            PDTResult(ch, pupils=[pa.CirclePupil(0.03), pa.CirclePupil(0.06)], max_size=n),
            TrackedPDTResult(ch, max_size=n)]


def _data(results):
    return [np.asarray(m.data, dtype=np.float64) for r in results for m in r.measures]


def test_fused_equals_generic(dev, monkeypatch):
    ch = _channel()
    np.random.seed(5)
    fused = _results(ch, 6)
    sim = Simulation(fused, batch=4)
    assert sim._fused_plan() is not None
    sim.run()
    np.random.seed(5)
    generic = _results(ch, 6)
    sim2 = Simulation(generic)
    monkeypatch.setattr(sim2, "_fused_plan", lambda: None)
    sim2.run()
    for a, b in zip(_data(fused), _data(generic)):
        assert a.shape == (6,)
        assert np.allclose(a, b, rtol=1e-6, atol=1e-12)


def test_fused_matches_oracle(dev):
    ch = _channel()
    np.random.seed(9)
    res = _results(ch, 3)
    Simulation(res, batch=3).run()
    rs = np.random.RandomState(9)
    ps = ch.path.phase_screens[0]
    for it in range(3):
        spectra = [o.draw_spectrum_numpy(rs, ps.f_grid.base, ps._get_psd()) for _ in range(3)]
        u = o.run_channel(o.gaussian_source(256, 0.002, 808e-9, 0.04, np.inf), 0.002, 808e-9, 5e3,
                          ch.path.positions, spectra)
        m = o.moments(u, 0.002)
        scale = m["mean_x2"]
        for j, key in enumerate(["mean_x", "mean_y", "mean_x2", "mean_xy", "mean_y2", "mean_x2_r"]):
            assert abs(res[0].measures[j].data[it] - m[key]) < 1e-4 * scale, key
        assert abs(res[1].measures[0].data[it] - o.pupil_eta(u, 0.002, 0.03)) < 1e-4
        assert abs(res[1].measures[1].data[it] - o.pupil_eta(u, 0.002, 0.06)) < 1e-4
        tracked = o.pupil_eta(u, 0.002, 0.05, (m["mean_x"], m["mean_y"]))
        assert abs(res[2].measures[2].data[it] - tracked) < 1e-4


def test_device_rng_reproducible(dev):
    ch = _channel()
    a = [BeamResult(ch, max_size=5)]
    Simulation(a, rng="device", seed=3, batch=5).run()
    b = [BeamResult(ch, max_size=5)]
    Simulation(b, rng="device", seed=3, batch=2).run()
    assert np.array_equal(np.asarray(a[0].measures[0].data), np.asarray(b[0].measures[0].data))
    assert np.all(np.isfinite(np.asarray(a[0].measures[2].data)))


def test_two_channels_draw_order(dev, monkeypatch):
    c1, c2 = _channel(), _channel(n=128, delta=0.004, S=2, P=64)
    np.random.seed(1)
    f = [BeamResult(c1, max_size=3), PDTResult(c2, max_size=3)]
    Simulation(f, batch=3).run()
    np.random.seed(1)
    g = [BeamResult(c1, max_size=3), PDTResult(c2, max_size=3)]
    s = Simulation(g)
    monkeypatch.setattr(s, "_fused_plan", lambda: None)
    s.run()
    for a, b in zip(_data(f), _data(g)):
        assert np.allclose(a, b, rtol=1e-6, atol=1e-12)


def test_wind_time_series_matches_oracle(dev):
    """Frozen-flow time series (simulations/wind.py): each realization's screens
    are drawn once (np.random legacy order) and re-evaluated shifted by (0, t)
    for every time; eta through the channel's pupil and the beam centroid at
    each shift against the FP64 oracle with the same draws and shifted screens."""
    from pyatmosphere_amd.simulations import TimeBWcorrSimulation, TimeCoherenceResult
    ch = _channel()
    times = [0.0, 0.3, 3.0]
    np.random.seed(13)
    tc = TimeCoherenceResult(ch, times, max_size=2)
    bw = TimeBWcorrSimulation(ch, times, max_size=2)
    Simulation([tc, bw]).run()
    rs = np.random.RandomState(13)
    ps = ch.path.phase_screens[0]
    etas = []
    for it in range(2):
        spectra = [o.draw_spectrum_numpy(rs, ps.f_grid.base, ps._get_psd()) for _ in range(3)]
        for ti, t in enumerate(times):
            u = o.run_channel(o.gaussian_source(256, 0.002, 808e-9, 0.04, np.inf), 0.002, 808e-9, 5e3,
                              ch.path.positions, spectra, shift=(0.0, t))
            m = o.moments(u, 0.002)
            etas.append(o.pupil_eta(u, 0.002, 0.05))
            assert abs(tc.measures[0].data[it][ti] - etas[-1]) < 1e-4
            assert abs(bw.measures[0].data[it][ti] - m["mean_x"]) < 1e-4 * m["mean_x2"]
            assert abs(bw.measures[1].data[it][ti] - m["mean_y"]) < 1e-4 * m["mean_y2"]
    assert etas[0] != etas[2]   # the shift changes what the pupil sees
    assert tc.tc[0] == pytest.approx(1.0)


def test_si_result_matches_oracle(dev):
    """SIResult (simulations/si.py): the on-axis intensity after every screen and
    at the receiver through the generic loop, against the FP64 oracle stepped
    screen by screen with the same np.random draws; the theory curve is the
    package's Andrews-Phillips on-axis SI at those positions."""
    from pyatmosphere_amd.simulations import SIResult
    from pyatmosphere_amd.theory.atmosphere.si import get_SI_andrews_strong
    ch = _channel()
    np.random.seed(21)
    r = SIResult(ch, max_size=3)
    Simulation([r]).run()
    got = r.intensities_at_center
    assert got.shape == (3, 4)
    rs = np.random.RandomState(21)
    ps = ch.path.phase_screens[0]
    ds = o.step_lengths(5e3, ch.path.positions)
    for it in range(3):
        spectra = [o.draw_spectrum_numpy(rs, ps.f_grid.base, ps._get_psd()) for _ in range(3)]
        u = o.gaussian_source(256, 0.002, 808e-9, 0.04, np.inf).astype(np.complex128)
        ref = []
        for k in range(3):
            u = np.exp(-1j * o.screen_phase(256, 0.002, spectra[k])) * o.vacuum_step(u, 0.002, 808e-9, ds[k])
            ref.append(abs(u[128, 128]) ** 2)
        u = o.vacuum_step(u, 0.002, 808e-9, ds[-1])
        ref.append(abs(u[128, 128]) ** 2)
        assert np.allclose(got[it], ref, rtol=1e-4, atol=0)
    assert np.allclose(r.theoretical_si[0], get_SI_andrews_strong(r.positions, ch.path.phase_screen.model, ch.source))
    assert r.si.shape == (4,)


@pytest.mark.parametrize("n", [64, 256, 2048, 4096])
def test_structure_function_kernel(dev, n):
    """pa_structure_function against the oracle's float64 restatement of
    theory/phase_screens/sf.py:4-8 averaged over rows, on random screens of
    size and spread like a composed screen (two at once; at 2048 and 4096 on
    a set of lags, the oracle's loop being slow), and on a linear ramp
    phi = c / 16, whose structure function is (r / 16)^2 (exactly up to 256:
    there the float32 sums of the squares are exact)."""
    import torch
    from pyatmosphere_amd import engine
    rng = np.random.default_rng(n)
    phi = (np.cumsum(rng.normal(size=(2, n, n)), axis=2) + 300.0).astype(np.float32)
    got = engine.structure_function(torch.from_numpy(phi).to(dev))
    assert got.shape == (2, n - 1)
    lags = np.arange(1, n) if n <= 256 else np.array([1, 2, 3, 17, 100, 1000, n // 2, n - 2, n - 1])
    p = phi.astype(np.float64)
    ref = np.stack([((p[..., :, :-r] - p[..., :, r:]) ** 2).sum(axis=-1).mean(axis=-1) / (n - r) for r in lags], -1)
    assert np.allclose(got[:, lags - 1], ref, rtol=2e-6, atol=0)
    ramp = np.tile(np.arange(n, dtype=np.float32) / 16, (n, 1))
    got = engine.structure_function(torch.from_numpy(ramp).to(dev))
    want = (np.arange(1, n) / 16.0) ** 2
    # larger n: each of the four float32 accumulators of a (row, lag) adds up
    # to n / 4 squares, each addition rounding by at most 2^-24 of the sum
    assert np.array_equal(got, want) if n <= 256 else np.allclose(got, want, rtol=1.5 * n / 4 * 2.0 ** -24, atol=0)


def test_structure_function_result_matches_oracle(dev):
    """StructureFunctionResult (simulations/structure_function.py): the first
    screen of each realization, device synthesis and device structure function,
    against the oracle's float64 screen from the same np.random draws."""
    from pyatmosphere_amd.simulations import StructureFunctionResult
    ch = _channel()
    np.random.seed(31)
    r = StructureFunctionResult(ch, max_size=2)
    Simulation([r]).run()
    got = np.asarray(r.measures[0].data)
    assert got.shape == (2, 255) and r.r.shape == (255,) and np.all(np.isfinite(r.get_numerical_theoretical))
    rs = np.random.RandomState(31)
    ps = ch.path.phase_screens[0]
    for it in range(2):
        spectra = [o.draw_spectrum_numpy(rs, ps.f_grid.base, ps._get_psd()) for _ in range(3)]
        ref = o.structure_function(o.screen_phase(256, 0.002, spectra[0]))
        assert np.allclose(got[it], ref, rtol=1e-3, atol=0)


def test_tracked_pupils_at_2048(dev):
    """Centroid-tracked transmittance on the 2048^2 / 50 km / 5-screen channel
    (8 device-RNG realizations, radii 0.1 and 0.2 m): the second reduction pass
    over the work buffer (StepRunner.tracked) equals the reduction of the
    stored fields with pupils centred on their own centroids
    (pa_field_partials with pup_moments), and the untracked centre pupil."""
    import torch
    from pyatmosphere_amd import engine
    from tests.helpers import channel_spec
    spec = channel_spec(n=2048, delta=0.0015, wvl=808e-9, w0=0.12, L=50e3, S=5, P=1024, Cn2=5e-16,
                        pupils=[(0.2, 0.0, 0.0)])
    B = 8
    tables = engine.DeviceTables([spec], dev)
    ws = engine.Workspace(2048, 1024, B, "f16x3", dev)
    field = torch.empty(B, 2048, 2048, dtype=torch.complex64, device=dev)
    mom, tr = engine.run_batch(ws, tables, np.zeros(B), np.full(B, 5), seed=11, rid=np.arange(B), field_out=field,
                               tracked=[0.1, 0.2])
    mom, tr = mom.cpu().numpy(), tr.cpu().numpy()
    ref = engine.field_moments(field, spec.delta, [(0.1, 0.0, 0.0), (0.2, 0.0, 0.0)], tracked=True)
    assert np.allclose(tr[:, 7:9], ref[:, 7:9], rtol=1e-9, atol=1e-14)
    assert np.allclose(tr[:, :7], mom[:, :7], rtol=1e-12, atol=0)
    # tracking moves the pupil: the transmittances differ from the centred pupil's
    assert np.any(np.abs(tr[:, 8] - mom[:, 7]) > 1e-6)
