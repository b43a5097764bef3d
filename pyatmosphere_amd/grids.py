"""Spatial and spectral grids (grids.py:7-119 of pyAtmosphere)."""
from dataclasses import dataclass

import numpy as np

from . import engine
from .gpu import device, get_xp


class Grid:
    def get_array_module(self):
        return get_xp()


@dataclass
class RectGrid(Grid):
    """Square-cell grid; coordinates (i + lb) * delta with lb = -N//2 + N%2
    (grids.py:13-85).  Device arrays are float32 like the reference's."""
    resolution: tuple
    delta: float

    def __post_init__(self):
        if isinstance(self.resolution, (int, np.integer)):
            self.resolution = (int(self.resolution), int(self.resolution))
        else:
            self.resolution = tuple(self.resolution)

    @property
    def size(self):
        return np.array(self.resolution) * self.delta

    @property
    def extent(self):
        return np.array([self._left_bound, self._right_bound, self._top_bound, self._bottom_bound]) * self.delta

    @property
    def shape(self):
        return self.resolution

    @property
    def origin_index(self):
        return (self.resolution[0] // 2, self.resolution[1] // 2)

    @property
    def _left_bound(self):
        return -self.resolution[0] // 2 + bool(self.resolution[0] % 2)

    @property
    def _right_bound(self):
        return self.resolution[0] // 2 + bool(self.resolution[0] % 2)

    @property
    def _top_bound(self):
        return -self.resolution[1] // 2 + bool(self.resolution[1] % 2)

    @property
    def _bottom_bound(self):
        return self.resolution[1] // 2 + bool(self.resolution[1] % 2)

    def _axis(self, lo, hi):
        xp = get_xp()
        return xp.arange(lo, hi, dtype=xp.float32, device=device()) * self.delta

    def get_NxNy(self):
        xp = get_xp()
        ny = xp.arange(self._top_bound, self._bottom_bound, device=device()).reshape(-1, 1)
        nx = xp.arange(self._left_bound, self._right_bound, device=device()).reshape(1, -1)
        return ny, nx

    def get_N2(self):
        a, b = self.get_NxNy()
        return a ** 2 + b ** 2

    def get_x(self):
        return self._axis(self._left_bound, self._right_bound).reshape(1, -1)

    def get_y(self):
        return self._axis(self._top_bound, self._bottom_bound).reshape(-1, 1)

    def get_xy(self):
        return self.get_x(), self.get_y()

    def get_rho2(self):
        x, y = self.get_xy()
        return x ** 2 + y ** 2

    def get_rho(self):
        return get_xp().sqrt(self.get_rho2())

    def get_f_grid(self):
        n = int(np.min(self.resolution))
        return RectGrid(resolution=n, delta=1 / (n * self.delta))

    # host float64 axis used by the device path
    def axis64(self):
        return engine.axis_coords(self.resolution[0], self.delta)


@dataclass
class RandLogPolarGrid(Grid):
    """Random log-polar spectral grid (grids.py:89-119).  Draws come from the
    host legacy numpy stream in the reference order: one radial jitter
    (np.random.random(1)) then the angles (np.random.random(P))."""
    points: int
    f_min: float
    f_max: float

    @property
    def base(self):
        return engine.log_polar_base(self.points, self.f_min, self.f_max)

    def draw_rho(self):
        rand = np.random.random(size=(1,)).astype(np.float32)
        f = self.base
        f_prev = np.insert(f, 0, 0)[:-1]
        return np.sqrt(f_prev ** 2 + rand * (f ** 2 - f_prev ** 2)).astype(np.float32)

    def draw_theta(self):
        return (2 * np.pi * np.random.random(size=(self.points,)).astype(np.float32)).astype(np.float32)

    def get_rho(self):
        return get_xp().from_numpy(self.draw_rho()).to(device())

    def get_theta(self):
        return get_xp().from_numpy(self.draw_theta()).to(device())

    def get_x(self, rho, theta):
        return rho * get_xp().cos(theta)

    def get_y(self, rho, theta):
        return rho * get_xp().sin(theta)

    def get_xy(self, rho, theta):
        xp = get_xp()
        return (rho * xp.cos(theta)).reshape((1, -1)), (rho * xp.sin(theta)).reshape((-1, 1))
