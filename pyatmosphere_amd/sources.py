"""Field sources (sources.py:8-23 of pyAtmosphere)."""
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr
from .gpu import device
from .theory.sources import GaussianBeam


@dataclass
class Source:
    wvl: float

    @property
    def k(self):
        return 2 * np.pi / self.wvl


class PlaneSource(Source):
    def output(self):
        return 1


class GaussianSource(GaussianBeam, Source):
    """Gaussian beam at z = 0 on the channel grid (sources.py:21-23;
    theory/sources.py:16-18), generated on the device from its two separable
    float64 factors."""

    def factors(self, n, delta):
        from .engine import gaussian_factors
        return gaussian_factors(n, delta, self.wvl, self.w0, self.F0)

    def output(self):
        grid = self.channel.grid
        n = grid.resolution[0]
        gy, gx = self.factors(n, grid.delta)
        dev = device()
        out = torch.empty(n, n, dtype=torch.complex64, device=dev)
        # keep the factor tensors referenced until the launch is enqueued: a freed
        # temporary's block can be handed to the next allocation before the kernel runs
        gy_d = torch.from_numpy(gy).to(dev)
        gx_d = torch.from_numpy(gx).to(dev)
        call("pa_outer", n, ptr(out), ptr(gy_d), ptr(gx_d), 1, _lib.stream_handle())
        return out
