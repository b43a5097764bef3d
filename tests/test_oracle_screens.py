"""Pins of the oracle's other screen generators (FFTPhaseScreen with
subharmonics, phase_screens.py:37-67; SUPhaseScreen, phase_screens.py:154-178)
on the host: the FFT form against a brute-force plane-wave sum, the ensemble
structure function against the MVK closed form, the SU amplitude law against
the SS annulus integrals, and the draw order of the package's host draws."""
import numpy as np
import pytest

import pyatmosphere_amd as pa
from oracle import fp64 as o
from pyatmosphere_amd.theory.models import MVKModel

K = 2 * np.pi / 808e-9


@pytest.mark.parametrize("sub", [0, 1, 3])
def test_fft_screen_equals_plane_wave_sum(sub):
    rs = np.random.RandomState(sub)
    co = o.draw_fft_coefficients(rs, 8, 0.01, K, 1e3, 1e-14, 5e-3, 10.0, sub)
    a, b = o.fft_screen(8, 0.01, co), o.fft_screen_bruteforce(8, 0.01, co)
    assert np.abs(a - b).max() < 1e-12 * np.abs(b).max()
    assert abs(a.mean()) < 1e-12 * np.abs(a).max()      # mean removed


def test_fft_screen_structure_function_near_mvk():
    """300 screens of 64^2 (delta 1 cm, Cn2 1e-14, l0 5 mm, L0 10 m, 3 levels):
    the mean row structure function is 0.85-0.95 of MVKModel.sf_phi for lags
    1-32 cells (the grid's Nyquist cut removes small-scale power, the
    subharmonics restore most of the large-scale power); a missing 2 pi df or
    square root in the amplitude would be off by factors."""
    n, d, R = 64, 0.01, 300
    rs = np.random.RandomState(1)
    D = np.zeros(n - 1)
    for _ in range(R):
        co = o.draw_fft_coefficients(rs, n, d, K, 1e3, 1e-14, 5e-3, 10.0, 3)
        D += o.structure_function(o.fft_screen(n, d, co).real)
    D /= R
    r = np.arange(1, n) * d
    ratio = D / MVKModel(1e-14, 5e-3, 10.0).sf_phi(r, K, 1e3)
    lags = [0, 1, 3, 7, 15, 31]
    assert np.all((ratio[lags] > 0.8) & (ratio[lags] < 1.0)), ratio[lags]


@pytest.mark.parametrize("u", [0.0, 0.5, 1.0])
def test_su_variance_matches_annulus_integrals(u):
    """Sum of the SU amplitude variances psd_phi_f(rho_k) pi delta_k_base over
    the points equals the SS screens' integrated annulus variances within 2%,
    for radial jitter 0, 1/2 and 1 (the spectrum sampled at the annulus inner
    edge, middle and outer edge)."""
    base = o.log_polar_base(1024, 1 / 15000, 2 / 6e-3).astype(np.float64)
    lo = np.concatenate([[0.0], base[:-1]])
    rho = np.sqrt(lo ** 2 + u * (base ** 2 - lo ** 2))
    su = o.psd_phi_f(rho, K, 1e4, 1e-15, 6e-3, 1e3) * np.pi * o.su_delta_k_base(base.astype(np.float32))
    ss = o.annulus_psd(base.astype(np.float32), K, 1e4, 1e-15, 6e-3, 1e3)
    assert abs(su.sum() / ss.sum() - 1) < 0.02


def _channel(screen, n=64, delta=0.01):
    ch = pa.Channel(grid=pa.RectGrid(resolution=n, delta=delta),
                    source=pa.GaussianSource(wvl=808e-9, w0=0.05, F0=np.inf),
                    path=pa.IdenticalPhaseScreensPath(phase_screen=screen, length=5e3, count=3))
    ch.path.init_phase_screens()
    return ch.path.phase_screens[0]


def test_package_su_draws_follow_the_reference_order():
    ps = _channel(pa.SUPhaseScreen(model=MVKModel(Cn2=1e-14, l0=5e-3, L0=10.0),
                                   f_grid=pa.RandLogPolarGrid(points=64, f_min=1 / 150, f_max=2 / 5e-3)))
    np.random.seed(4)
    sp = ps._get_spectrum()
    ref = o.draw_su_spectrum(np.random.RandomState(4), ps.f_grid.base, K, ps.thickness, 1e-14, 5e-3, 10.0)
    assert np.array_equal(sp.rho, ref.rho) and np.array_equal(sp.theta, ref.theta)
    assert np.array_equal(sp.value, ref.value)


def test_package_fft_screen_host_algebra():
    """The package's FFT-screen draws equal the oracle's, and the device
    algorithm restated on the host (signed unnormalised inverse DFT, the
    subharmonics as separable float64 terms, the analytic mean) equals the
    oracle's screen."""
    n = 64
    ps = _channel(pa.FFTPhaseScreen(3, model=MVKModel(Cn2=1e-14, l0=5e-3, L0=10.0)), n=n)
    np.random.seed(4)
    main, subs = ps.draw()
    ey, g, mean = ps._separable(subs)
    co = o.draw_fft_coefficients(np.random.RandomState(4), n, 0.01, K, ps.thickness, 1e-14, 5e-3, 10.0, 3)
    assert np.array_equal(main, co.main)
    x = np.arange(n)
    sign = (-1.0) ** (x[:, None] + x[None, :])
    host = sign * np.fft.ifft2(sign * main.astype(np.complex128)) * n * n + (ey[:, :, None] * g[:, None, :]).sum(0)
    ref = o.fft_screen(n, 0.01, co)
    assert np.abs(host - mean - ref).max() < 1e-10 * np.abs(ref).max()
