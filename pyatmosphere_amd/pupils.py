"""Circular aperture (pupils.py:4-13 of pyAtmosphere)."""
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import call, ptr


@dataclass
class CirclePupil:
    radius: float

    def get_pupil(self, shift=(0, 0)):
        """Boolean device mask (x - sx)^2 + (y + sy)^2 <= radius^2."""
        x, y = self.channel.grid.get_xy()
        return (x.double() - shift[0]) ** 2 + (y.double() + shift[1]) ** 2 <= self.radius ** 2

    def output(self, input, shift=(0, 0), **kwargs):
        """input * mask as a new device array."""
        grid = self.channel.grid
        u = input.to(torch.complex64).clone().contiguous()
        n = u.shape[-1]
        batch = u.numel() // (n * n)
        call("pa_apply_pupil", n, ptr(u), float(grid.delta), float(self.radius), float(shift[0]),
             float(shift[1]), batch, _lib.stream_handle())
        return u
