"""FP64 CPU oracle for the split-step phase-screen path.

TEST INFRASTRUCTURE ONLY. This module restates, in float64/complex128 numpy, the
mathematics of pyAtmosphere's split-step hot path so that the HIP implementation
in ``pyatmosphere_amd`` can be checked against it. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import it,
and only as the checker. The product package never imports anything from here.

Every function cites the reference file:line it restates (paths are relative to
the pyAtmosphere source tree).

Statistical judge: ``r_bw`` (Klyatskin-Kon beam wander) is pinned by the
reference notebook's own simulated sigma_BW,x (main.ipynb:252, within 5%).
The other theory closed forms (Rytov variance, r0, the MVK structure function,
the Andrews-Phillips scintillation index) are not restated here: the package's
own versions are pinned directly against independent limits in
tests/test_theory.py.

Conventions fixed here (and followed by the device path):
  * grid coordinates are exact float64 ``(i + lb) * delta`` with
    ``lb = -N//2 + N%2`` (grids.py:40-53, 63-69), rows are y, columns are x;
  * spectral coefficient records are float32 ``rho``, float32 ``theta`` and
    complex64 ``value`` exactly as the reference stores them
    (grids.py:98-107, phase_screens.py:98-103); the oracle upcasts them;
  * the centred FFT pair with delta**2 and (N*df)**2 scales (utils.py:42-50)
    reduces exactly to ``ifft2(H(fftfreq) * fft2(u))``; the oracle keeps the
    shifted form in ``vacuum_step_shifted`` so a test can show the equality.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

try:        # the same pocketfft algorithm as numpy.fft, with threads (OMP_NUM_THREADS when set)
    import os as _os
    import scipy.fft as _fft
    _FFT_KW = {"workers": int(_os.environ.get("OMP_NUM_THREADS", 0)) or _os.cpu_count() or 1}
except ImportError:  # pragma: no cover
    _fft = np.fft
    _FFT_KW = {}


# ---------------------------------------------------------------- grids
def axis_coords(n: int, delta: float) -> np.ndarray:
    """RectGrid.get_x / get_y coordinates (grids.py:40-53, 63-69), float64."""
    lb = -n // 2 + (n % 2)
    return (np.arange(n, dtype=np.float64) + lb) * float(delta)


def grid_xy(n: int, delta: float) -> Tuple[np.ndarray, np.ndarray]:
    """x as a (1, N) row and y as an (N, 1) column (grids.py:63-72)."""
    a = axis_coords(n, delta)
    return a.reshape(1, -1), a.reshape(-1, 1)


def f_grid_delta(n: int, delta: float) -> float:
    """RectGrid.get_f_grid spacing 1/(N*delta) (grids.py:82-85)."""
    return 1.0 / (n * delta)


# ---------------------------------------------------------------- source
def gaussian_source(n: int, delta: float, wvl: float, w0: float, F0: float) -> np.ndarray:
    """GaussianBeam.amplitude on RectGrid.get_rho2 (theory/sources.py:16-18;
    sources.py:21-23; grids.py:74-76)."""
    x, y = grid_xy(n, delta)
    r2 = x ** 2 + y ** 2
    curv = 0.0 if np.isinf(F0) else 2 * np.pi / wvl / 2 / F0
    return np.sqrt(2 / np.pi) / w0 * np.exp(-(1 / w0 ** 2 + 1j * curv) * r2)


def gaussian_w(length: float, wvl: float, w0: float, F0: float) -> float:
    """Closed-form beam radius w(z) (theory/sources.py:20-33)."""
    k = 2 * np.pi / wvl
    theta0 = 1 - length / F0
    lam0 = 2 * length / k / w0 ** 2
    return w0 * np.sqrt(theta0 ** 2 + lam0 ** 2)


# ---------------------------------------------------------------- vacuum step
def transfer_phase_1d(n: int, delta: float, wvl: float, length: float) -> np.ndarray:
    """pi*lambda*d*f^2 on the unshifted FFT frequency axis, float64
    (theory/vacuum.py:7 with the f-grid of grids.py:82-85)."""
    f = np.fft.fftfreq(n, d=delta)
    return np.pi * length * wvl * f ** 2


def vacuum_step(u: np.ndarray, delta: float, wvl: float, length: float) -> np.ndarray:
    """Angular-spectrum Fresnel step ``ifft2(e^{ikd} e^{-i pi lambda d f^2} fft2(u))``
    (theory/vacuum.py:5-7; pathes.py:27-40). ``length <= 0`` is the identity
    (pathes.py:30, 39-40)."""
    if not length > 0:
        return u
    n = u.shape[0]
    k = 2 * np.pi / wvl
    ph = transfer_phase_1d(n, delta, wvl, length)
    H = np.exp(1j * k * length) * np.exp(-1j * (ph[:, None] + ph[None, :]))
    return _fft.ifft2(H * _fft.fft2(u, **_FFT_KW), **_FFT_KW)


def vacuum_step_shifted(u: np.ndarray, delta: float, wvl: float, length: float) -> np.ndarray:
    """The same step written with the reference's centred transforms and scales
    (utils.py:42-50; theory/vacuum.py:7): used to show that the shifts cancel."""
    n = u.shape[0]
    k = 2 * np.pi / wvl
    df = f_grid_delta(n, delta)
    fx, fy = grid_xy(n, df)
    f2 = fx ** 2 + fy ** 2
    U = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(u))) * delta ** 2
    U = np.exp(1j * k * length) * np.exp(-1j * np.pi * length * (2 * np.pi / k) * f2) * U
    return np.fft.ifftshift(np.fft.ifft2(np.fft.ifftshift(U))) * (n * df) ** 2


# ---------------------------------------------------------------- turbulence model
def mvk_psd_n(kappa, Cn2: float, l0: float, L0: float):
    """MVKModel.psd_n (theory/models.py:80-85)."""
    k0 = 2 * np.pi / L0
    km = 5.92 / l0
    return 0.033 * Cn2 * np.exp(-(kappa / km) ** 2) / (kappa ** 2 + k0 ** 2) ** (11 / 6)


def psd_phi_f(f, k: float, thickness: float, Cn2: float, l0: float, L0: float):
    """Model.psd_phi_f (theory/models.py:16-23)."""
    return 2 * np.pi * k ** 2 * thickness * mvk_psd_n(2 * np.pi * f, Cn2, l0, L0)


def log_polar_base(points: int, f_min: float, f_max: float) -> np.ndarray:
    """RandLogPolarGrid.base in float32 (grids.py:94-96)."""
    return np.exp(np.linspace(np.log(f_min), np.log(f_max), points, dtype=np.float32))


_GL_X, _GL_W = np.polynomial.legendre.leggauss(48)


def annulus_psd(base: np.ndarray, k: float, thickness: float, Cn2: float, l0: float,
                L0: float, panels: int = 4) -> np.ndarray:
    """Per-annulus phase variance ``2 pi * int_{f_{i-1}}^{f_i} (2 pi)^2 f Phi(f) df``
    with ``f_{-1} = 0`` (phase_screens.py:82-91), by composite Gauss-Legendre
    (48 nodes x ``panels`` panels per annulus) in float64, returned as float64."""
    f = base.astype(np.float64)
    lo = np.concatenate([[0.0], f[:-1]])
    hi = f
    out = np.zeros_like(f)
    edges = np.linspace(0.0, 1.0, panels + 1)
    for p in range(panels):
        a = lo + (hi - lo) * edges[p]
        b = lo + (hi - lo) * edges[p + 1]
        mid = 0.5 * (a + b)[:, None]
        half = 0.5 * (b - a)[:, None]
        ff = mid + half * _GL_X[None, :]
        vals = (2 * np.pi) ** 2 * ff * psd_phi_f(ff, k, thickness, Cn2, l0, L0)
        out += (half[:, 0]) * (vals @ _GL_W)
    return 2 * np.pi * out


# ---------------------------------------------------------------- spectra
@dataclass
class Spectrum:
    """One screen's coefficient record, stored as the reference stores it
    (utils.py:35-39; grids.py:98-107; phase_screens.py:98-103)."""
    rho: np.ndarray     # float32 (P,)
    theta: np.ndarray   # float32 (P,)
    value: np.ndarray   # complex64 (P,)


def draw_spectrum_numpy(rs: np.random.RandomState, base: np.ndarray, psd32: np.ndarray) -> Spectrum:
    """Draws in the reference order random(1) -> random(P) -> normal((2, P))
    (phase_screens.py:98-103; grids.py:98-107) from a legacy RandomState."""
    P = base.shape[0]
    rand = rs.random_sample(size=(1,)).astype(np.float32)
    f = base
    f_prev = np.insert(f, 0, 0)[:-1]
    rho = np.sqrt(f_prev ** 2 + rand * (f ** 2 - f_prev ** 2)).astype(np.float32)
    theta = (2 * np.pi * rs.random_sample(size=(P,)).astype(np.float32)).astype(np.float32)
    nrm = rs.normal(size=(2, P))
    value = ((nrm[0] + 1j * nrm[1]).astype(np.complex64) * np.sqrt(psd32)).astype(np.complex64)
    return Spectrum(rho=rho, theta=theta, value=value)


def screen_phase(n: int, delta: float, spec: Spectrum, shift=(0.0, 0.0)) -> np.ndarray:
    """SS screen ``Re[(a * E_y) @ E_x]`` in complex128 (phase_screens.py:108-126, 25-28;
    grids.py:109-119). Returns float64 (N, N)."""
    rho = spec.rho.astype(np.float64)
    th = spec.theta.astype(np.float64)
    fx = rho * np.cos(th)
    fy = rho * np.sin(th)
    x, y = grid_xy(n, delta)
    x = x + shift[0]
    y = y + shift[1]
    Ey = np.exp(1j * 2 * np.pi * (y @ fy.reshape(1, -1)))        # (N, P)
    Ex = np.exp(1j * 2 * np.pi * (fx.reshape(-1, 1) @ x))        # (P, N)
    a = spec.value.astype(np.complex128)
    A = a[None, :] * Ey
    # the real part of the complex product, as two real float64 products
    # (contiguous copies of the parts, so that both go to BLAS)
    c = np.ascontiguousarray
    return c(A.real) @ c(Ex.real) - c(A.imag) @ c(Ex.imag)


def screen_phase_bruteforce(n: int, delta: float, spec: Spectrum, shift=(0.0, 0.0)) -> np.ndarray:
    """Direct plane-wave double sum for tiny sizes (phase_screens.py:125-126)."""
    rho = spec.rho.astype(np.float64)
    th = spec.theta.astype(np.float64)
    fx = rho * np.cos(th)
    fy = rho * np.sin(th)
    xs = axis_coords(n, delta) + shift[0]
    ys = axis_coords(n, delta) + shift[1]
    out = np.zeros((n, n))
    for r in range(n):
        for c in range(n):
            s = 0j
            for k in range(len(fx)):
                s += complex(spec.value[k]) * np.exp(1j * 2 * np.pi * (fx[k] * xs[c] + fy[k] * ys[r]))
            out[r, c] = s.real
    return out


# ---------------------------------------------------------------- path
def path_positions(length: float, count: int, position_in_slab: str = "middle") -> np.ndarray:
    """IdenticalPhaseScreensPath screen positions (pathes.py:79-89)."""
    thickness = length / count
    if position_in_slab == "before":
        return np.arange(count) * thickness
    if position_in_slab == "middle":
        return (np.arange(count) + 1 / 2) * thickness
    if position_in_slab == "after":
        return (np.arange(count) + 1) * thickness
    raise ValueError(position_in_slab)


def step_lengths(length: float, positions: Sequence[float]) -> List[float]:
    """Fresnel step lengths of PhaseScreensPath.generator (pathes.py:67-75)."""
    d = []
    for i in range(len(positions)):
        d.append(positions[i] - positions[i - 1] if i > 0 else positions[0])
    d.append(length - positions[-1])
    return [float(v) for v in d]


def append_losses(u, losses_db: float, path_losses_db: float):
    """AbstractPath.append_losses (pathes.py:19-21): the amplitude factor
    10^(-dB/20) with dB = ``losses_db or path_losses_db`` (so a zero-dB segment
    takes the whole path's loss), and no factor when that is zero."""
    db = losses_db or path_losses_db
    return u * 10 ** (-db / 20) if db else u


def run_channel(u0: np.ndarray, delta: float, wvl: float, length: float,
                positions: Sequence[float], spectra: Sequence[Spectrum],
                shift=(0.0, 0.0), return_screens: bool = False, losses_db: float = 0.0,
                total_losses: bool = False, return_steps: bool = False):
    """Split-step loop ``u <- losses_i(exp(-i phi_i) P(d_i) u)`` then ``P(d_S)``
    (pathes.py:61-75).  Segment i loses ``losses_db d_i / length`` dB through
    append_losses (pathes.py:71-73); the final step none (its VacuumPath has no
    loss).  ``total_losses``: AbstractPath.output applies the whole path's loss
    once more (pathes.py:23-24).  ``return_steps``: also the field after each
    screen (what the generator yields)."""
    n = u0.shape[0]
    u = u0.astype(np.complex128)
    ds = step_lengths(length, positions)
    screens, steps = [], []
    for i, spec in enumerate(spectra):
        phi = screen_phase(n, delta, spec, shift)
        screens.append(phi)
        u = append_losses(np.exp(-1j * phi) * vacuum_step(u, delta, wvl, ds[i]),
                          losses_db * ds[i] / length, losses_db)
        steps.append(u)
    u = vacuum_step(u, delta, wvl, ds[-1])
    if total_losses:
        u = append_losses(u, None, losses_db)
    out = (u,)
    if return_screens:
        out += (screens,)
    if return_steps:
        out += (steps,)
    return out if len(out) > 1 else u


# ---------------------------------------------------------------- pupil and measures
def pupil_mask(n: int, delta: float, radius: float, shift=(0.0, 0.0)) -> np.ndarray:
    """CirclePupil.get_pupil (pupils.py:8-10)."""
    x, y = grid_xy(n, delta)
    return (x - shift[0]) ** 2 + (y + shift[1]) ** 2 <= radius ** 2


def moments(u: np.ndarray, delta: float) -> dict:
    """measures.eta/mean_x/mean_y/mean_x2/mean_xy/mean_y2 on a given field
    (measures.py:1-38) and BeamResult.mean_x2_r (simulations/beam.py:26-33)."""
    n = u.shape[0]
    x, y = grid_xy(n, delta)
    I = np.abs(u) ** 2
    d2 = delta ** 2
    out = {
        "eta": float(I.sum() * d2),
        "mean_x": float((I * x).sum() * d2),
        "mean_y": float((I * (-1) * y).sum() * d2),
        "mean_x2": float((I * x ** 2).sum() * d2),
        "mean_xy": float((I * (-1 * x * y)).sum() * d2),
        "mean_y2": float((I * y ** 2).sum() * d2),
    }
    r0 = np.sqrt(out["mean_x"] ** 2 + out["mean_y"] ** 2)
    cx = out["mean_x"] / r0
    sx = out["mean_y"] / r0
    out["mean_x2_r"] = float((I * (x * cx + (-1 * y) * sx) ** 2).sum() * d2)
    return out


def pupil_eta(u: np.ndarray, delta: float, radius: float, shift=(0.0, 0.0)) -> float:
    """PDTResult.append_pupil followed by measures.eta (simulations/pdt.py:21-26;
    measures.py:7-8)."""
    m = pupil_mask(u.shape[0], delta, radius, shift)
    return float((np.abs(u * m) ** 2).sum() * delta ** 2)


def structure_function(phi: np.ndarray) -> np.ndarray:
    """Row structure function averaged over rows, in float64: calculate_sf
    (theory/phase_screens/sf.py:4-8) then .mean(axis=1)
    (simulations/structure_function.py:11-12).  phi (n, n) -> (n - 1,);
    (B, n, n) -> (B, n - 1)."""
    p = np.asarray(phi, dtype=np.float64)
    n = p.shape[-1]
    d = np.stack([((p[..., :, :-r] - p[..., :, r:]) ** 2).sum(axis=-1).mean(axis=-1) / (n - r)
                  for r in range(1, n)], axis=-1)
    return d


# ---------------------------------------------------------------- theory judges
def rytov2(Cn2: float, k: float, length: float) -> float:
    """get_rytov2 (theory/atmosphere/__init__.py:5-6)."""
    return 1.23 * Cn2 * k ** (7 / 6) * length ** (11 / 6)


def r_bw(L: float, Cn2: float, l0: float, L0: float, wvl: float, w0: float, F0: float) -> float:
    """get_r_bw, Klyatskin-Kon beam wander (theory/atmosphere/beam_wandering.py:5-15)."""
    import scipy.integrate
    k = 2 * np.pi / wvl

    def dsp(rho):
        kappa0 = 2 * np.pi / L0
        return 1.09 * Cn2 * k ** 2 * L * l0 ** (-1 / 3) * rho ** 2 * (
            1 / (1 + (rho / l0) ** 2) ** (1 / 6) - 0.72 * (kappa0 * l0) ** (1 / 3))

    def drc2(z, kappa):
        w = gaussian_w(z, wvl, w0, F0)
        return 4 * np.pi ** 2 * (L - z) ** 2 * kappa ** 3 * mvk_psd_n(kappa, Cn2, l0, L0) * np.exp(
            -(kappa * w) ** 2 / 4 - np.pi * dsp(kappa * z / k))

    return float(np.sqrt(scipy.integrate.dblquad(drc2, 0, np.inf, 0, L)[0]))


# ---------------------------------------------------------------- other screen generators
def su_delta_k_base(base: np.ndarray) -> np.ndarray:
    """SUPhaseScreen.delta_k_base (phase_screens.py:158-163): (2 pi)^2 (f_k^2 - f_{k-1}^2),
    f_{-1} = 0, in float64."""
    f = base.astype(np.float64)
    return (2 * np.pi) ** 2 * (f ** 2 - np.concatenate([[0.0], f[:-1]]) ** 2)


def draw_su_spectrum(rs: np.random.RandomState, base: np.ndarray, k: float, thickness: float, Cn2: float,
                     l0: float, L0: float) -> Spectrum:
    """SUPhaseScreen.generate_phase_screen's draws (phase_screens.py:165-178): the
    radii (one radial jitter, grids.py:98-103) and angles (grids.py:105-107) as
    for the SS screen, then normal((2, P)); amplitude
    sqrt(psd_phi_f(rho) pi delta_k_base) at the drawn radius, in float64,
    stored as complex64."""
    P = base.shape[0]
    rand = rs.random_sample(size=(1,)).astype(np.float32)
    f_prev = np.insert(base, 0, 0)[:-1]
    rho = np.sqrt(f_prev ** 2 + rand * (base ** 2 - f_prev ** 2)).astype(np.float32)
    theta = (2 * np.pi * rs.random_sample(size=(P,)).astype(np.float32)).astype(np.float32)
    nrm = rs.normal(size=(2, P))
    amp = np.sqrt(psd_phi_f(rho.astype(np.float64), k, thickness, Cn2, l0, L0) * np.pi * su_delta_k_base(base))
    value = ((nrm[0] + 1j * nrm[1]) * amp).astype(np.complex64)
    return Spectrum(rho=rho, theta=theta, value=value)


def screen_complex(n: int, delta: float, spec: Spectrum, shift=(0.0, 0.0)) -> np.ndarray:
    """The complex plane-wave sum (a * E_y) @ E_x itself (phase_screens.py:125-126
    and 178): generate(complex=True) of the SS and SU screens."""
    rho = spec.rho.astype(np.float64)
    th = spec.theta.astype(np.float64)
    fx, fy = rho * np.cos(th), rho * np.sin(th)
    x, y = grid_xy(n, delta)
    Ey = np.exp(1j * 2 * np.pi * ((y + shift[1]) @ fy.reshape(1, -1)))
    Ex = np.exp(1j * 2 * np.pi * (fx.reshape(-1, 1) @ (x + shift[0])))
    return (spec.value.astype(np.complex128)[None, :] * Ey) @ Ex


@dataclass
class FFTCoefficients:
    """One FFT screen's draws, stored as the reference stores them: the N x N
    complex64 coefficients on the centred f-grid and the 3 x 3 coefficients of
    each subharmonic level (phase_screens.py:42-64)."""
    main: np.ndarray                     # (N, N) complex64
    sub: List[np.ndarray]                # per level (3, 3) complex64
    sub_delta: List[float]               # per level frequency spacing


def _cn(rs, res: int, df: float, k: float, thickness: float, Cn2: float, l0: float, L0: float) -> np.ndarray:
    """get_cn_coefficients (phase_screens.py:45-50) on RectGrid(res, df): complex
    Gaussian draws normal(shape) + 1j normal(shape) times
    sqrt(psd_phi_f(|f|)) 2 pi df, zero at the origin index (res // 2, res // 2);
    |f| from the float32 grid (grids.py:63-80), the rest in float64."""
    lb = -res // 2 + res % 2
    ax = np.arange(lb, lb + res, dtype=np.float32) * np.float32(df)
    rho = np.sqrt(ax.reshape(1, -1) ** 2 + ax.reshape(-1, 1) ** 2)
    nrm = (rs.normal(size=(res, res)) + 1j * rs.normal(size=(res, res))).astype(np.complex64)
    cn = (nrm.astype(np.complex128) * np.sqrt(psd_phi_f(rho.astype(np.float64), k, thickness, Cn2, l0, L0))
          * 2 * np.pi * df).astype(np.complex64)
    cn[res // 2, res // 2] = 0
    return cn


def draw_fft_coefficients(rs: np.random.RandomState, n: int, delta: float, k: float, thickness: float,
                          Cn2: float, l0: float, L0: float, subharmonics: int) -> FFTCoefficients:
    """FFTPhaseScreen.generate_phase_screen's draws in the reference order: the
    f-grid RectGrid(n, 1/(n delta)) first, then each subharmonic level sh on
    RectGrid(3, df / 3^(sh + 1)) (phase_screens.py:52-57)."""
    df = 1.0 / (n * delta)
    main = _cn(rs, n, df, k, thickness, Cn2, l0, L0)
    sub, sd = [], []
    for sh in range(subharmonics):
        d = df / 3 ** (sh + 1)
        sub.append(_cn(rs, 3, d, k, thickness, Cn2, l0, L0))
        sd.append(d)
    return FFTCoefficients(main=main, sub=sub, sub_delta=sd)


def fft_screen(n: int, delta: float, co: FFTCoefficients) -> np.ndarray:
    """FFTPhaseScreen.generate_phase_screen (phase_screens.py:42-67) in complex128:
    ifft2(cn, 1) (utils.py:47-50: the centred inverse transform times n^2), plus
    cn_ij exp(2 pi i (f_i x + f_j y)) for each subharmonic level (i indexes the
    x frequency), minus the mean over the grid."""
    c = co.main.astype(np.complex128)
    phase = np.fft.ifftshift(np.fft.ifft2(np.fft.ifftshift(c))) * n ** 2
    x, y = grid_xy(n, delta)
    for cn, d in zip(co.sub, co.sub_delta):
        f = np.array([-1.0, 0.0, 1.0]) * d
        for i in range(3):
            for j in range(3):
                phase = phase + complex(cn[i, j]) * np.exp(1j * 2 * np.pi * (f[i] * x + f[j] * y))
    return phase - phase.mean()


def fft_screen_bruteforce(n: int, delta: float, co: FFTCoefficients) -> np.ndarray:
    """The same screen as one direct plane-wave sum over every f-grid point and
    subharmonic (tiny n only)."""
    xs = axis_coords(n, delta)
    df = 1.0 / (n * delta)
    fa = axis_coords(n, df)
    waves = [(fa[c], fa[r], complex(co.main[r, c])) for r in range(n) for c in range(n)]
    for cn, d in zip(co.sub, co.sub_delta):
        f = np.array([-1.0, 0.0, 1.0]) * d
        waves += [(f[i], f[j], complex(cn[i, j])) for i in range(3) for j in range(3)]
    out = np.zeros((n, n), np.complex128)
    for fx, fy, a in waves:
        out += a * np.exp(1j * 2 * np.pi * (fx * xs[None, :] + fy * xs[:, None]))
    return out - out.mean()
