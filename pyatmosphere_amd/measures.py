"""Intensity-weighted observables (measures.py:1-38 of pyAtmosphere).

Each measure reduces |u|^2 on the device (one pass computing every moment in
float64) and returns a Python float.  Without `output` the channel is run, as
in the reference.
"""
import torch

from . import _lib, engine
from ._lib import call, ptr


def I(channel, output=None, *args, **kwargs):
    u = output if output is not None else channel.run(*args, **kwargs)
    u = u.to(torch.complex64).contiguous()
    out = torch.empty(u.shape, dtype=torch.float32, device=u.device)
    call("pa_intensity", u.numel(), ptr(u), ptr(out), _lib.stream_handle())
    return out


def _moments(channel, output=None, *args, **kwargs):
    u = output if output is not None else channel.run(*args, **kwargs)
    return engine.field_moments(u, channel.grid.delta)[0]


def eta(channel, *args, **kwargs):
    return float(_moments(channel, *args, **kwargs)[0])


def mean_x(channel, *args, **kwargs):
    kwargs["pupil"] = False
    return float(_moments(channel, *args, **kwargs)[1])


def mean_y(channel, *args, **kwargs):
    kwargs["pupil"] = False
    return float(_moments(channel, *args, **kwargs)[2])


def mean_x2(channel, *args, **kwargs):
    kwargs["pupil"] = False
    return float(_moments(channel, *args, **kwargs)[3])


def mean_xy(channel, *args, **kwargs):
    kwargs["pupil"] = False
    return float(_moments(channel, *args, **kwargs)[4])


def mean_y2(channel, *args, **kwargs):
    kwargs["pupil"] = False
    return float(_moments(channel, *args, **kwargs)[5])
