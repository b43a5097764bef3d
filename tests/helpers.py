"""Shared builders for the parity tests: a channel description used by both the
device engine and the FP64 oracle, and seeded host spectra."""
import numpy as np

from oracle import fp64 as o
from pyatmosphere_amd import engine


def channel_spec(n, delta, wvl, w0, L, S, P, Cn2, l0=6e-3, L0=1e3, f_min=1 / 15000, f_max=2 / 6e-3,
                 pupils=(), F0=np.inf, position="middle"):
    base = engine.log_polar_base(P, f_min, f_max)
    k = 2 * np.pi / wvl
    psd = engine.annulus_psd(base, k, L / S, lambda kap: engine.mvk_psd_n(kap, Cn2, l0, L0)).astype(np.float32)
    return engine.ChannelSpec(n=n, delta=delta, wvl=wvl, w0=w0, F0=F0, length=L,
                              positions=o.path_positions(L, S, position), points=P, base=base, psd=psd,
                              pupils=list(pupils))


def host_spectra(spec, batch, seed):
    """Reference-order draws (np.random legacy stream) for `batch` realizations:
    returns per screen (rho, theta, value) arrays of shape (batch, P) and the
    per-realization oracle spectra lists."""
    rs = np.random.RandomState(seed)
    per_real = [[o.draw_spectrum_numpy(rs, spec.base, spec.psd) for _ in range(spec.nscreens)]
                for _ in range(batch)]
    screens = []
    for i in range(spec.nscreens):
        screens.append((np.stack([per_real[b][i].rho for b in range(batch)]),
                        np.stack([per_real[b][i].theta for b in range(batch)]),
                        np.stack([per_real[b][i].value for b in range(batch)])))
    return screens, per_real


def oracle_field(spec, spectra):
    u0 = o.gaussian_source(spec.n, spec.delta, spec.wvl, spec.w0, spec.F0)
    return o.run_channel(u0, spec.delta, spec.wvl, spec.length, spec.positions, spectra, shift=spec.shift)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
