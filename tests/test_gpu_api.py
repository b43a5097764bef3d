"""pyAtmosphere-compatible API on the device: the reference notebook's vacuum
known answers and the turbulent path against the oracle with the same seed."""
import numpy as np
import pytest
import torch

import pyatmosphere_amd as pa
from oracle import fp64 as o
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture
def vacuum_channel(dev):
    # tests/itest_vacuum_propagation.ipynb cell 3
    return pa.Channel(grid=pa.RectGrid(resolution=2 ** 8, delta=0.002),
                      source=pa.GaussianSource(wvl=809e-9, w0=0.02, F0=np.inf),
                      path=pa.VacuumPath(length=4e3), pupil=pa.CirclePupil(radius=0.5))


def test_vacuum_full_eta(vacuum_channel):
    np.testing.assert_almost_equal(pa.measures.eta(vacuum_channel, pupil=False), 1, decimal=6)


def test_vacuum_norm_amplitude(vacuum_channel):
    u = vacuum_channel.run(pupil=False)
    actual = abs(u.to(torch.complex128).sum().item()) * vacuum_channel.grid.delta ** 2
    np.testing.assert_almost_equal(actual, np.sqrt(2 * np.pi) * vacuum_channel.source.w0, decimal=7)


def test_vacuum_norm_intensity(vacuum_channel):
    I = pa.measures.I(vacuum_channel, pupil=False)
    np.testing.assert_almost_equal(I.double().sum().item() * vacuum_channel.grid.delta ** 2, 1, decimal=5)


def test_vacuum_waist(vacuum_channel):
    ch = vacuum_channel
    zR = np.pi * ch.source.w0 ** 2 / ch.source.wvl
    desired = ch.source.w0 * np.sqrt(1 + (ch.path.length / zR) ** 2)
    mean_r2 = pa.measures.mean_x2(ch) + pa.measures.mean_y2(ch)
    np.testing.assert_almost_equal(np.sqrt(2 * mean_r2), desired, decimal=7)


def _quick(n=256, delta=0.002, L=5e3, S=3):
    return pa.Channel(
        grid=pa.RectGrid(resolution=n, delta=delta),
        source=pa.GaussianSource(wvl=808e-9, w0=0.05, F0=np.inf),
        path=pa.IdenticalPhaseScreensPath(
            phase_screen=pa.SSPhaseScreen(model=pa.MVKModel(Cn2=1e-15, l0=6e-3, L0=1e3),
                                          f_grid=pa.RandLogPolarGrid(points=128, f_min=1 / 15000, f_max=2 / 6e-3)),
            length=L, count=S),
        pupil=pa.CirclePupil(radius=0.06))


def _oracle_for(ch, seed):
    rs = np.random.RandomState(seed)
    ps = ch.path.phase_screens[0]
    spectra = [o.draw_spectrum_numpy(rs, ps.f_grid.base, ps._get_psd()) for _ in ch.path.phase_screens]
    u0 = o.gaussian_source(ch.grid.resolution[0], ch.grid.delta, ch.source.wvl, ch.source.w0, ch.source.F0)
    return o.run_channel(u0, ch.grid.delta, ch.source.wvl, ch.path.length, ch.path.positions, spectra)


def test_run_matches_oracle_same_seed(dev):
    ch = _quick()
    np.random.seed(0)
    u = ch.run(pupil=False)
    ref = _oracle_for(ch, 0)
    assert rel_l2(u.cpu().numpy(), ref) < 1e-4
    np.random.seed(0)
    up = ch.run()
    mask = o.pupil_mask(256, ch.grid.delta, 0.06)
    assert rel_l2(up.cpu().numpy(), ref * mask) < 1e-4
    assert np.isclose(pa.measures.eta(ch, output=up), o.pupil_eta(ref, ch.grid.delta, 0.06), rtol=1e-4)


def test_generator_matches_run(dev):
    ch = _quick()
    np.random.seed(3)
    ref = ch.run(pupil=False)
    np.random.seed(3)
    steps = list(ch.generator(pupil=False))
    assert len(steps) == 3
    assert all(u.shape == (256, 256) and phi.shape == (256, 256) for u, phi in steps)
    assert torch.equal(ch.output, ref)
    # the field after each screen (pathes.py:61-75) against the oracle stepped
    # screen by screen with the same draws
    rs = np.random.RandomState(3)
    ps = ch.path.phase_screens[0]
    spectra = [o.draw_spectrum_numpy(rs, ps.f_grid.base, ps._get_psd()) for _ in range(3)]
    u = o.gaussian_source(256, 0.002, 808e-9, 0.05, np.inf).astype(np.complex128)
    for k, d in enumerate(o.step_lengths(5e3, ch.path.positions)[:3]):
        u = np.exp(-1j * o.screen_phase(256, 0.002, spectra[k])) * o.vacuum_step(u, 0.002, 808e-9, d)
        assert rel_l2(steps[k][0].cpu().numpy(), u) < 1e-4


def test_screen_generate_matches_oracle(dev):
    ch = _quick()
    ps = ch.path.phase_screens[0]
    ch.path.init_phase_screens()
    np.random.seed(7)
    phi = ps.generate()
    rs = np.random.RandomState(7)
    ref = o.screen_phase(256, ch.grid.delta, o.draw_spectrum_numpy(rs, ps.f_grid.base, ps._get_psd()))
    assert np.abs(phi.cpu().numpy() - ref).max() < 1e-4 + 2e-7 * np.abs(ref).max()


def test_measures_match_oracle(dev):
    ch = _quick()
    np.random.seed(1)
    u = ch.run(pupil=False)
    m = o.moments(u.cpu().numpy().astype(np.complex128), ch.grid.delta)
    for name in ["mean_x", "mean_y", "mean_x2", "mean_xy", "mean_y2"]:
        assert np.isclose(getattr(pa.measures, name)(ch, output=u), m[name], rtol=1e-9, atol=1e-15)


def test_use_gpu_false_raises(dev):
    pa.gpu.config["use_gpu"] = False
    try:
        with pytest.raises(RuntimeError, match="GPU only"):
            _quick().run()
    finally:
        pa.gpu.config["use_gpu"] = True
