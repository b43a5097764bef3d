"""Component wiring helpers and the spectral record (utils.py:7-39 of pyAtmosphere)."""
from dataclasses import dataclass
from typing import Sequence


class CrossRef:
    """Attribute that also sets a back reference on the assigned component
    (utils.py:7-21): `channel.grid = g` makes `g.channel` the channel."""

    def __init__(self, cross_ref_name):
        self.cross_ref_name = cross_ref_name

    def __set_name__(self, owner, name):
        self.private_name = "_" + name

    def __get__(self, obj, objtype=None):
        if obj is None:
            return self
        return getattr(obj, self.private_name, None)

    def __set__(self, obj, value):
        if not value:
            return
        setattr(obj, self.private_name, value)
        setattr(value, self.cross_ref_name, obj)


class Default:
    """Read-through default resolved along a dotted attribute path
    (utils.py:24-32), e.g. Default("channel.source.wvl")."""

    def __init__(self, default_path):
        self._path = default_path.split(".")

    def __get__(self, obj, objtype=None):
        if obj is None:
            return self
        node = obj
        for name in self._path:
            node = getattr(node, name)
        return node


@dataclass
class PolarDiscreteFunction:
    """One screen's spectral draw (utils.py:35-39): radii, angles and complex
    amplitudes.  Held on the host (P values each)."""
    rho: Sequence[float]
    theta: Sequence[float]
    value: Sequence[float]
