"""The other screen generators on the device against the FP64 oracle with the
same host draws (np.random seed): SUPhaseScreen (phase_screens.py:154-178),
FFTPhaseScreen with subharmonics (phase_screens.py:37-67), the complex
screens of generate(complex=True) and PhaseScreen.generator's real/imaginary
pairs (phase_screens.py:25-34), and channels built on SU and FFT screens."""
import numpy as np
import pytest
import torch

import pyatmosphere_amd as pa
from oracle import fp64 as o
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

N, DELTA, WVL, CN2, L0_IN, L0_OUT = 256, 0.002, 808e-9, 1e-15, 6e-3, 1e3


def _channel(screen, L=5e3, S=3, n=N, delta=DELTA):
    return pa.Channel(grid=pa.RectGrid(resolution=n, delta=delta),
                      source=pa.GaussianSource(wvl=WVL, w0=0.05, F0=np.inf),
                      path=pa.IdenticalPhaseScreensPath(phase_screen=screen, length=L, count=S),
                      pupil=pa.CirclePupil(radius=0.06))


def _su(points=128):
    return pa.SUPhaseScreen(model=pa.MVKModel(Cn2=CN2, l0=L0_IN, L0=L0_OUT),
                            f_grid=pa.RandLogPolarGrid(points=points, f_min=1 / 15000, f_max=2 / 6e-3))


def _fft(sub=3):
    return pa.FFTPhaseScreen(sub, model=pa.MVKModel(Cn2=CN2, l0=L0_IN, L0=L0_OUT))


def _screen_tol(ref):
    # the float32 screen plus the split-fp16 remainder: 1e-4 rad + float32 rounding of |phi|
    return 1e-4 + 2e-7 * np.abs(ref).max()


def test_su_screen_matches_oracle(dev):
    ch = _channel(_su())
    ps = ch.path.phase_screens[0]
    ch.path.init_phase_screens()
    np.random.seed(5)
    phi = ps.generate(complex=True)
    rs = np.random.RandomState(5)
    sp = o.draw_su_spectrum(rs, ps.f_grid.base, 2 * np.pi / WVL, ps.thickness, CN2, L0_IN, L0_OUT)
    ref = o.screen_complex(N, DELTA, sp)
    got = phi.cpu().numpy()
    assert phi.dtype == torch.complex64
    assert np.abs(got.real - ref.real).max() < _screen_tol(ref.real)
    assert np.abs(got.imag - ref.imag).max() < _screen_tol(ref.imag)


def test_ss_generator_yields_real_and_imaginary(dev):
    ch = _channel(pa.SSPhaseScreen(model=pa.MVKModel(Cn2=CN2, l0=L0_IN, L0=L0_OUT),
                                   f_grid=pa.RandLogPolarGrid(points=128, f_min=1 / 15000, f_max=2 / 6e-3)))
    ps = ch.path.phase_screens[0]
    ch.path.init_phase_screens()
    np.random.seed(9)
    gen = ps.generator()
    re, im = next(gen), next(gen)
    rs = np.random.RandomState(9)
    ref = o.screen_complex(N, DELTA, o.draw_spectrum_numpy(rs, ps.f_grid.base, ps._get_psd()))
    assert np.abs(re.cpu().numpy() - ref.real).max() < _screen_tol(ref.real)
    assert np.abs(im.cpu().numpy() - ref.imag).max() < _screen_tol(ref.imag)


@pytest.mark.parametrize("n,delta,sub", [(256, 0.002, 3), (2048, 1.5e-3, 3), (512, 0.004, 0)])
def test_fft_screen_matches_oracle(dev, n, delta, sub):
    ch = _channel(_fft(sub), n=n, delta=delta)
    ps = ch.path.phase_screens[0]
    ch.path.init_phase_screens()
    np.random.seed(17)
    phi = ps.generate(complex=True).cpu().numpy()
    rs = np.random.RandomState(17)
    co = o.draw_fft_coefficients(rs, n, delta, 2 * np.pi / WVL, ps.thickness, CN2, L0_IN, L0_OUT, sub)
    ref = o.fft_screen(n, delta, co)
    scale = np.abs(ref).max()
    print(f"fft screen n={n}: max |phi| {scale:.3g}, max error re {np.abs(phi.real - ref.real).max():.2e}",
          flush=True)
    # float32 inverse FFT: ~1e-6 of the screen's scale
    assert np.abs(phi.real - ref.real).max() < 1e-5 * scale + 1e-6
    assert np.abs(phi.imag - ref.imag).max() < 1e-5 * scale + 1e-6


@pytest.mark.parametrize("make", [_su, _fft])
def test_channel_with_other_screens_matches_oracle(dev, make):
    """Channel.run and the generator over SU or FFT screens against the oracle's
    split-step loop with the same screens."""
    ch = _channel(make())
    ps0 = ch.path.phase_screens[0]
    ch.path.init_phase_screens()
    k = 2 * np.pi / WVL
    np.random.seed(23)
    u = ch.run(pupil=False).cpu().numpy()
    rs = np.random.RandomState(23)
    screens = []
    for ps in ch.path.phase_screens:
        if isinstance(ps0, pa.SUPhaseScreen):
            sp = o.draw_su_spectrum(rs, ps.f_grid.base, k, ps.thickness, CN2, L0_IN, L0_OUT)
            screens.append(o.screen_complex(N, DELTA, sp).real)
        else:
            co = o.draw_fft_coefficients(rs, N, DELTA, k, ps.thickness, CN2, L0_IN, L0_OUT, ps.subharmonics)
            screens.append(o.fft_screen(N, DELTA, co).real)
    ref = o.gaussian_source(N, DELTA, WVL, 0.05, np.inf).astype(np.complex128)
    ds = o.step_lengths(ch.path.length, ch.path.positions)
    steps = []
    for i, phi in enumerate(screens):
        ref = np.exp(-1j * phi) * o.vacuum_step(ref, DELTA, WVL, ds[i])
        steps.append(ref)
    ref = o.vacuum_step(ref, DELTA, WVL, ds[-1])
    assert rel_l2(u, ref) < 1e-4
    np.random.seed(23)
    gen = list(ch.generator(pupil=False))
    for (uu, phi), r, s in zip(gen, steps, screens):
        assert rel_l2(uu.cpu().numpy(), r) < 1e-4
        assert np.abs(phi.cpu().numpy() - s).max() < _screen_tol(s)
    assert rel_l2(ch.output.cpu().numpy(), ref) < 1e-4
