"""Phase screens (phase_screens.py:9-179 of pyAtmosphere).

SSPhaseScreen and SUPhaseScreen keep the reference's interface: spectral draws
in the reference order from the host legacy numpy stream, and `generate`
returning the screen as a device array (float32, or complex64 with
complex=True).  Both are plane-wave sums over random log-polar points
synthesised by libpa_hip.so (plane-wave tables with float64 argument
reduction, float64 1-D parts, split-fp16 MFMA GEMM for the 2-D remainder); they
differ only in the amplitude law.  FFTPhaseScreen draws its coefficients on
the centred f-grid and its subharmonic levels on the host (reference order)
and runs the centred inverse FFT and the subharmonic sum on the device
(pa_fft_screen).
"""
from typing import Tuple

import numpy as np
import torch

from . import engine
from .gpu import config, device
from .utils import Default, PolarDiscreteFunction


class PhaseScreen:
    wvl = Default("channel.source.wvl")
    grid = Default("channel.grid")

    def __init__(self, model, thickness=None, wvl=None, grid=None):
        self.model = model
        self.thickness = thickness
        if wvl:
            self.wvl = wvl
        if grid:
            self.grid = grid

    def generate_phase_screen(self, *args, **kwargs):
        """Return the screen (device array)."""
        raise NotImplementedError

    def generate(self, complex=False, *args, **kwargs):
        """The screen (phase_screens.py:25-28): complex64 with complex=True,
        else its real part (float32)."""
        return self.generate_phase_screen(*args, complex=complex, **kwargs)

    def generator(self, *args, **kwargs):
        """Two screens per synthesis: the real and the imaginary part of one
        complex screen (phase_screens.py:30-34)."""
        while True:
            ps = self.generate(True, *args, **kwargs)
            yield ps.real
            yield ps.imag


def _synthesize(screen, spectrum: PolarDiscreteFunction, shift=(0.0, 0.0), imaginary: bool = False) -> torch.Tensor:
    """Re sum a_k exp(i 2 pi (fx_k x + fy_k y)) (phase_screens.py:118-126) of one
    spectral record on `screen`'s grid as a float32 device array; with
    `imaginary`, the imaginary part instead (the real part of the sum with
    amplitudes -i a_k)."""
    n = screen.grid.resolution[0]
    P = len(spectrum.rho)
    dev = device()
    ws = engine.get_workspace(n, P, 1, config.get("gemm_mode", "f16x3"), dev)
    spec = engine.ChannelSpec(n=n, delta=screen.grid.delta, wvl=screen.wvl, w0=1.0, F0=np.inf, length=0.0,
                              positions=np.zeros(1), points=P, base=np.asarray(spectrum.rho, np.float32),
                              psd=np.zeros(P, np.float32), shift=tuple(float(v) for v in shift))
    tables = engine.DeviceTables([spec], dev)
    value = np.asarray(spectrum.value, np.complex64)
    if imaginary:
        value = (value * np.complex64(-1j)).astype(np.complex64)
    engine.upload_spectra(ws, np.asarray(spectrum.rho)[None], np.asarray(spectrum.theta)[None], value[None], batch=1)
    ws.chan[:1].zero_()
    stream = engine._lib.stream_handle()
    p2 = ws.phi2_buffer()
    engine.synth_screen(ws, tables, 0, 1, None, stream, phi2=p2)
    phi = torch.empty(1, n, n, dtype=torch.float32, device=dev)
    engine.call("pa_screen_compose", n, engine.ptr(ws.sparam), engine.ptr(ws.hg), engine.ptr(p2),
                engine.ptr(phi), 1, stream)
    return phi[0]


def _screen_of(screen, spectrum, shift, complex):
    re = _synthesize(screen, spectrum, shift)
    if not complex:
        return re
    return torch.complex(re, _synthesize(screen, spectrum, shift, imaginary=True))


class SSPhaseScreen(PhaseScreen):
    def __init__(self, f_grid, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.f_grid = f_grid
        self._psd = None
        self.cache_clear()

    def cache_clear(self):
        self._cached_spectrum = None

    def _get_psd(self) -> np.ndarray:
        """float32 (P,) annulus variances, cached (phase_screens.py:82-91)."""
        if self._psd is not None:
            return self._psd
        k = 2 * np.pi / self.wvl
        self._psd = engine.annulus_psd(self.f_grid.base, k, self.thickness, self.model.psd_n).astype(np.float32)
        return self._psd

    def _get_spectrum(self, use_cached_spectrum=False) -> PolarDiscreteFunction:
        """Draw order random(1) -> random(P) -> normal((2, P)) (phase_screens.py:93-106)."""
        if use_cached_spectrum and self._cached_spectrum is not None:
            return self._cached_spectrum
        rho = self.f_grid.draw_rho()
        theta = self.f_grid.draw_theta()
        nrm = np.random.normal(size=(2, self.f_grid.points))
        value = ((nrm[0] + 1j * nrm[1]).astype(np.complex64) * np.sqrt(self._get_psd())).astype(np.complex64)
        spectrum = PolarDiscreteFunction(rho=rho, theta=theta, value=value)
        if use_cached_spectrum:
            self._cached_spectrum = spectrum
        return spectrum

    def synthesize(self, spectrum: PolarDiscreteFunction, shift=(0.0, 0.0)) -> torch.Tensor:
        """Full screen Re sum a_k exp(i 2 pi (fx_k x + fy_k y)) as a float32 device array."""
        return _synthesize(self, spectrum, shift)

    def generate_phase_screen(self, shift: Tuple[float, float] = (0, 0), wind: bool = False, complex=False):
        spectrum = self._get_spectrum(use_cached_spectrum=wind)
        return _screen_of(self, spectrum, shift, complex)


class SUPhaseScreen(PhaseScreen):
    """Random log-polar plane-wave screen whose amplitudes sample the phase
    spectrum at the drawn radius (phase_screens.py:154-178):
    a_k = (n + i m) sqrt(Phi_phi(rho_k) pi delta_k_base[k]),
    delta_k_base = (2 pi)^2 (f_k^2 - f_{k-1}^2).  Draws: radial jitter, angles,
    normal((2, P)) from the host legacy numpy stream."""

    def __init__(self, f_grid, *args, **kwargs):
        self.f_grid = f_grid
        super().__init__(*args, **kwargs)
        self._delta_k_base = None
        self.cache_clear()

    def cache_clear(self):
        self._cached_spectrum = None

    @property
    def delta_k_base(self) -> np.ndarray:
        if self._delta_k_base is None:
            f = self.f_grid.base.astype(np.float64)
            self._delta_k_base = (2 * np.pi) ** 2 * (f ** 2 - np.concatenate([[0.0], f[:-1]]) ** 2)
        return self._delta_k_base

    def _get_psd(self) -> np.ndarray:
        """Placeholder annulus variances: the SU amplitudes depend on each draw's
        radii, so device-side draws (which scale by this) are not used for SU
        screens; the host draws carry the amplitudes."""
        return np.zeros(self.f_grid.points, np.float32)

    def _get_spectrum(self, use_cached_spectrum=False) -> PolarDiscreteFunction:
        if use_cached_spectrum and self._cached_spectrum is not None:
            return self._cached_spectrum
        rho = self.f_grid.draw_rho()
        theta = self.f_grid.draw_theta()
        nrm = np.random.normal(size=(2, self.f_grid.points))
        k = 2 * np.pi / self.wvl
        amp = np.sqrt(self.model.psd_phi_f(rho.astype(np.float64), k, self.thickness) * np.pi * self.delta_k_base)
        spectrum = PolarDiscreteFunction(rho=rho, theta=theta, value=((nrm[0] + 1j * nrm[1]) * amp).astype(np.complex64))
        if use_cached_spectrum:
            self._cached_spectrum = spectrum
        return spectrum

    def generate_phase_screen(self, shift: Tuple[float, float] = (0, 0), wind: bool = False, complex=False):
        return _screen_of(self, self._get_spectrum(use_cached_spectrum=wind), shift, complex)


class FFTPhaseScreen(PhaseScreen):
    """Screen from random coefficients on the centred f-grid of the channel's
    grid, inverse-FFT'd, plus `subharmonics` levels of 3 x 3 low-frequency
    plane waves, minus its mean (phase_screens.py:37-67).  The coefficients are
    drawn on the host in the reference order (normal(N, N) + 1j normal(N, N),
    then each level's normal(3, 3) pair)."""

    def __init__(self, subharmonics, *args, **kwargs):
        self.subharmonics = subharmonics
        super().__init__(*args, **kwargs)
        self._scale = None

    def cache_clear(self):
        pass

    def _amplitude(self, res: int, df: float) -> np.ndarray:
        """sqrt(psd_phi_f(|f|)) 2 pi df on RectGrid(res, df) (float32 |f|, as the
        reference's grid; float64 otherwise)."""
        lb = -res // 2 + res % 2
        ax = np.arange(lb, lb + res, dtype=np.float32) * np.float32(df)
        rho = np.sqrt(ax.reshape(1, -1) ** 2 + ax.reshape(-1, 1) ** 2).astype(np.float64)
        return np.sqrt(self.model.psd_phi_f(rho, 2 * np.pi / self.wvl, self.thickness)) * 2 * np.pi * df

    def _coefficients(self, res: int, df: float, scale=None) -> np.ndarray:
        """get_cn_coefficients (phase_screens.py:45-50)."""
        nrm = (np.random.normal(size=(res, res)) + 1j * np.random.normal(size=(res, res))).astype(np.complex64)
        a = scale if scale is not None else self._amplitude(res, df)
        cn = (nrm.astype(np.complex128) * a).astype(np.complex64)
        cn[res // 2, res // 2] = 0
        return cn

    def draw(self):
        """(main coefficients (N, N) complex64, [(level coefficients (3, 3), level spacing)])."""
        n = self.grid.resolution[0]
        df = 1.0 / (n * self.grid.delta)
        if self._scale is None or self._scale[0] != (n, df, self.wvl, self.thickness):
            self._scale = ((n, df, self.wvl, self.thickness), self._amplitude(n, df))
        main = self._coefficients(n, df, self._scale[1])
        subs = []
        for sh in range(self.subharmonics):
            d = df / 3 ** (sh + 1)
            subs.append((self._coefficients(3, d), d))
        return main, subs

    def _separable(self, subs):
        """Host float64 factors of the subharmonic sum and the screen's mean:
        sum_{i,j} cn_ij e^{2 pi i f_i x} e^{2 pi i f_j y} = sum_j ey_j(r) g_j(c)."""
        n = self.grid.resolution[0]
        xs = engine.axis_coords(n, self.grid.delta)
        ey, g = [], []
        for cn, d in subs:
            f = np.array([-1.0, 0.0, 1.0]) * d
            ex = np.exp(1j * 2 * np.pi * f[:, None] * xs[None, :])        # (3, n): e^{2 pi i f_i x_c}
            for j in range(3):
                ey.append(ex[j])                                         # rows use the same axis
                g.append((cn[:, j].astype(np.complex128)[:, None] * ex).sum(axis=0))
        ey = np.array(ey).reshape(-1, n)
        g = np.array(g).reshape(-1, n)
        mean = complex((ey.mean(axis=1) * g.mean(axis=1)).sum()) if len(ey) else 0j
        return ey, g, mean

    def screen(self, complex=False, angle_out=None):
        """One screen on the device: (N, N) float32 (complex64 with complex=True);
        with `angle_out` (1, N, N) float32, also its real part in the work
        layout (what the split-step row pass applies)."""
        n = self.grid.resolution[0]
        dev = device()
        main, subs = self.draw()
        ey, g, mean = self._separable(subs)
        coeff = torch.from_numpy(main).to(dev).reshape(1, n, n)
        work = torch.empty(1, n, n, dtype=torch.complex64, device=dev)
        nsub = ey.shape[0]
        sub = torch.from_numpy(np.concatenate([ey, g]).astype(np.complex128)).to(dev) if nsub else None
        mean_t = torch.tensor([mean], dtype=torch.complex128, device=dev)
        re = torch.empty(1, n, n, dtype=torch.float32, device=dev)
        im = torch.empty(1, n, n, dtype=torch.float32, device=dev) if complex else None
        engine.call("pa_fft_screen", n, engine.ptr(coeff), engine.ptr(work), engine.ptr(sub), nsub, engine.ptr(mean_t),
                    engine.ptr(engine.twiddles(n, dev)), engine.ptr(re), engine.ptr(im), engine.ptr(angle_out), 1,
                    engine._lib.stream_handle())
        return torch.complex(re[0], im[0]) if complex else re[0]

    def generate_phase_screen(self, shift: Tuple[float, float] = (0, 0), wind: bool = False, complex=False):
        if tuple(shift) != (0, 0):
            raise NotImplementedError("FFTPhaseScreen has no shifted (wind) form (phase_screens.py:42-67)")
        return self.screen(complex=complex)
