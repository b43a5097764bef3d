"""Sparse-spectral phase screens (phase_screens.py:9-136 of pyAtmosphere).

SSPhaseScreen keeps the reference's interface: a lazily computed per-annulus
variance (`_get_psd`), a spectral draw in the reference order from the host
legacy numpy stream (`_get_spectrum`), and `generate` returning the float32
screen as a device array.  The screen itself is synthesised by libpa_hip.so
(plane-wave tables with float64 argument reduction, float64 1-D parts, split
fp16 MFMA GEMM for the 2-D remainder).
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
        if complex:
            raise NotImplementedError("the device synthesis forms the real screen only")
        return self.generate_phase_screen(*args, **kwargs)

    def generator(self, *args, **kwargs):
        while True:
            yield self.generate(*args, **kwargs)


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
        n = self.grid.resolution[0]
        P = self.f_grid.points
        dev = device()
        ws = engine.get_workspace(n, P, 1, config.get("gemm_mode", "f16x3"), dev)
        spec = engine.ChannelSpec(n=n, delta=self.grid.delta, wvl=self.wvl, w0=1.0, F0=np.inf, length=0.0,
                                  positions=np.zeros(1), points=P, base=self.f_grid.base,
                                  psd=self._get_psd(), shift=tuple(float(v) for v in shift))
        tables = engine.DeviceTables([spec], dev)
        engine.upload_spectra(ws, np.asarray(spectrum.rho)[None], np.asarray(spectrum.theta)[None],
                              np.asarray(spectrum.value)[None], batch=1)
        ws.chan[:1].zero_()
        stream = engine._lib.stream_handle()
        p2 = ws.phi2_buffer()
        engine.synth_screen(ws, tables, 0, 1, None, stream, phi2=p2)
        phi = torch.empty(1, n, n, dtype=torch.float32, device=dev)
        engine.call("pa_screen_compose", n, engine.ptr(ws.sparam), engine.ptr(ws.hg), engine.ptr(p2),
                    engine.ptr(phi), 1, stream)
        return phi[0]

    def generate_phase_screen(self, shift: Tuple[float, float] = (0, 0), wind: bool = False):
        spectrum = self._get_spectrum(use_cached_spectrum=wind)
        return self.synthesize(spectrum, shift)
