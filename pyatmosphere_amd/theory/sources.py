"""Gaussian beam closed forms (theory/sources.py:6-33 of pyAtmosphere)."""
import numpy as np


class GaussianBeam:
    def __init__(self, wvl, w0, F0):
        self.wvl = wvl
        self.w0 = w0
        self.F0 = F0

    @property
    def k(self):
        return 2 * np.pi / self.wvl

    def get_theta0(self, length):
        return 1 - length / self.F0

    def get_Lambda0(self, length):
        return 2 * length / self.k / self.w0 ** 2

    def get_theta(self, length):
        t0, l0 = self.get_theta0(length), self.get_Lambda0(length)
        return t0 / (t0 ** 2 + l0 ** 2)

    def get_Lambda(self, length):
        t0, l0 = self.get_theta0(length), self.get_Lambda0(length)
        return l0 / (t0 ** 2 + l0 ** 2)

    def get_w(self, length):
        """Beam radius w(z) = w0 sqrt(Theta0^2 + Lambda0^2)."""
        return self.w0 * np.sqrt(self.get_theta0(length) ** 2 + self.get_Lambda0(length) ** 2)
