"""Propagation paths (pathes.py:10-95 of pyAtmosphere) on the device.

VacuumPath runs one fused Fresnel step; PhaseScreensPath runs the whole
split-step loop through libpa_hip.so (row/column FFT passes with the transfer
function and the screen multiply fused in), drawing each screen's spectrum
from its SSPhaseScreen in the reference order.
"""
import copy
from abc import ABC, abstractmethod

import numpy as np
import torch

from . import engine
from .gpu import config, device
from .phase_screens import FFTPhaseScreen, SSPhaseScreen, SUPhaseScreen

_PLANE_WAVE = (SSPhaseScreen, SUPhaseScreen)


def _as_field(input, n, dev):
    """The input field as a (1, n, n) complex64 device tensor: a tensor or array
    of n x n values, or a scalar broadcast over the grid."""
    if torch.is_tensor(input):
        u = input.to(device=dev, dtype=torch.complex64)
    else:
        a = np.asarray(input)
        if a.ndim == 0:
            u = torch.full((n, n), complex(a), dtype=torch.complex64, device=dev)
        else:
            u = torch.as_tensor(a.astype(np.complex64), device=dev)
    if u.numel() != n * n:
        raise ValueError(f"input field has {u.numel()} values, the grid {n} x {n}")
    return u.reshape(1, n, n).contiguous()


def _gain(losses_db):
    return 10 ** (-losses_db / 20) if losses_db else 1.0


class AbstractPath(ABC):
    def __init__(self, length, losses_db=0):
        self.length = length
        self.losses_db = losses_db

    @abstractmethod
    def lossless_output(self, input, *args, **kwargs):
        pass

    def append_losses(self, input, losses_db=None):
        """input * 10^(-losses/20) (pathes.py:19-21, including `losses_db or self.losses_db`)."""
        losses_db = losses_db or self.losses_db
        return input * 10 ** (-losses_db / 20) if losses_db else input

    def output(self, input, *args, **kwargs):
        return self.append_losses(self.lossless_output(input, *args, **kwargs))


class VacuumPath(AbstractPath):
    def lossless_output(self, input, length=None):
        length = length if length is not None else self.length
        if not length > 0:
            return input
        return _vacuum(self.channel, input, length)


def _vacuum(channel, input, length, gain=1.0):
    grid = channel.grid
    n = grid.resolution[0]
    dev = device()
    spec = engine.ChannelSpec(n=n, delta=grid.delta, wvl=channel.source.wvl, w0=1.0, F0=np.inf,
                              length=float(length), positions=np.zeros(0), points=16,
                              base=np.ones(16, np.float32), psd=np.zeros(16, np.float32), gains=[gain])
    tables = engine.DeviceTables([spec], dev)
    ws = engine.get_workspace(n, 16, 1, config.get("gemm_mode", "f16x3"), dev)
    out = torch.empty(1, n, n, dtype=torch.complex64, device=dev)
    engine.run_batch(ws, tables, np.zeros(1), np.zeros(1), field_in=_as_field(input, n, dev),
                     field_out=out, reduce=False)
    return out[0]


class PhaseScreensPath(AbstractPath):
    def __init__(self, length, phase_screens, positions, losses_db=0):
        self.positions = positions
        self.phase_screens = phase_screens
        super().__init__(length, losses_db)

    def init_phase_screens(self):
        for phase_screen in self.phase_screens:
            phase_screen.channel = self.channel

    # -- device plumbing
    def _segment_gains(self, with_total):
        ds = engine.ChannelSpec.step_lengths(self)  # uses self.positions / self.length
        gains = []
        for i in range(len(self.positions)):
            part = self.losses_db * ds[i] / self.length
            gains.append(_gain(part or self.losses_db))
        gains.append(_gain(self.losses_db) if with_total else 1.0)
        return gains

    def _spec(self, shift, with_total):
        self.init_phase_screens()
        screens = self.phase_screens
        for ps in screens:
            if not isinstance(ps, _PLANE_WAVE + (FFTPhaseScreen,)):
                raise NotImplementedError("the device path runs SS, SU and FFT phase screens")
        waves = [ps for ps in screens if isinstance(ps, _PLANE_WAVE)]
        if waves:
            P = waves[0].f_grid.points
            if any(ps.f_grid.points != P for ps in waves):
                raise NotImplementedError("all plane-wave screens of a path must use the same number of points")
            base, psd = waves[0].f_grid.base, waves[0]._get_psd()
        else:   # FFT screens only: no spectral points to synthesise
            P, base, psd = 16, np.ones(16, np.float32), np.zeros(16, np.float32)
        grid = self.channel.grid
        src = self.channel.source
        return engine.ChannelSpec(
            n=grid.resolution[0], delta=grid.delta, wvl=src.wvl, w0=getattr(src, "w0", 1.0),
            F0=getattr(src, "F0", np.inf), length=float(self.length), positions=np.asarray(self.positions, float),
            points=P, base=base, psd=psd, shift=(float(shift[0]), float(shift[1])),
            gains=self._segment_gains(with_total))

    def _draw(self, wind):
        out = []
        for ps in self.phase_screens:
            sp = ps._get_spectrum(use_cached_spectrum=wind)
            out.append((np.asarray(sp.rho)[None], np.asarray(sp.theta)[None], np.asarray(sp.value)[None]))
        return out

    def _screen_step(self, run, i, ps, wind, compose, shift=(0, 0)):
        """Screen i of a step-level run: plane-wave screens are synthesised from
        their host draw into the runner's workspace; FFT screens write their
        screen angle there themselves.  Applies the screen; returns the real
        screen (row-major float32) when `compose`."""
        if isinstance(ps, _PLANE_WAVE):
            sp = ps._get_spectrum(use_cached_spectrum=wind)
            return run.screen(i, host=(np.asarray(sp.rho)[None], np.asarray(sp.theta)[None],
                                       np.asarray(sp.value)[None]), compose=compose)
        if tuple(shift) != (0, 0):
            raise NotImplementedError("FFTPhaseScreen has no shifted (wind) form (phase_screens.py:42-67)")
        phi = ps.screen(angle_out=run.ws.ang[:1])
        run.apply(i)
        return phi[None] if compose else None

    def _run(self, input, shift=(0, 0), wind=False, with_total=False):
        spec = self._spec(shift, with_total)
        dev = device()
        tables = engine.DeviceTables([spec], dev)
        ws = engine.get_workspace(spec.n, spec.points, 1, config.get("gemm_mode", "f16x3"), dev)
        out = torch.empty(1, spec.n, spec.n, dtype=torch.complex64, device=dev)
        if all(isinstance(ps, _PLANE_WAVE) for ps in self.phase_screens):
            engine.run_batch(ws, tables, np.zeros(1), np.array([spec.nscreens]), host_spectra=self._draw(wind),
                             field_in=_as_field(input, spec.n, dev), field_out=out, reduce=False)
            return out[0]
        run = engine.StepRunner(ws, tables, np.zeros(1), np.array([spec.nscreens]))
        run.begin(_as_field(input, spec.n, dev))
        for i, ps in enumerate(self.phase_screens):
            run.fresnel(i)
            self._screen_step(run, i, ps, wind, compose=False, shift=shift)
        run.fresnel(len(self.phase_screens))
        run.finish(out, reduce=False)
        return out[0]

    def lossless_output(self, input, shift=(0, 0), wind=False, complex=False):
        return self._run(input, shift, wind, with_total=False)

    def output(self, input, shift=(0, 0), wind=False, complex=False):
        return self._run(input, shift, wind, with_total=True)

    def generator(self, input, shift=(0, 0), wind=False, complex=False):
        """Yields (field after screen i, screen i) and returns the output field
        (pathes.py:61-75)."""
        spec = self._spec(shift, with_total=False)
        dev = device()
        tables = engine.DeviceTables([spec], dev)
        ws = engine.get_workspace(spec.n, spec.points, 1, config.get("gemm_mode", "f16x3"), dev)
        run = engine.StepRunner(ws, tables, np.zeros(1), np.array([spec.nscreens]))
        run.begin(_as_field(input, spec.n, dev))
        for i, ps in enumerate(self.phase_screens):
            run.fresnel(i)
            phi = self._screen_step(run, i, ps, wind, compose=True, shift=shift)
            u = torch.empty(1, spec.n, spec.n, dtype=torch.complex64, device=dev)
            run.finish(u, reduce=False)
            # the work field after a screen is the unnormalised forward row
            # transform of the next step; the inverse row transform returns n u
            # (the transfer function carries the 1/n^2 of each full step), so
            # scale by 1/n (a power of two: exact)
            u.mul_(1.0 / spec.n)
            yield u[0], phi[0]
        run.fresnel(len(self.phase_screens))
        out = torch.empty(1, spec.n, spec.n, dtype=torch.complex64, device=dev)
        run.finish(out, reduce=False)
        return out[0]


class IdenticalPhaseScreensPath(PhaseScreensPath):
    def __init__(self, length, count, phase_screen, position_in_slab="middle", losses_db=0):
        thickness = length / count
        if position_in_slab == "before":
            positions = np.arange(count) * thickness
        elif position_in_slab == "middle":
            positions = (np.arange(count) + 1 / 2) * thickness
        elif position_in_slab == "after":
            positions = (np.arange(count) + 1) * thickness
        else:
            raise ValueError("Available values for position_in_slab: 'before', 'middle' and 'after'")
        phase_screen.thickness = thickness
        phase_screens = [copy.copy(phase_screen) for _ in range(count)]
        self.phase_screen = phase_screens[0]
        super().__init__(length=length, phase_screens=phase_screens, positions=positions, losses_db=losses_db)
