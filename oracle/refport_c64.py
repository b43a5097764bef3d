"""Op-for-op numpy restatement of the reference's host path, used as the CPU
baseline that bench.py times (the reference itself does not travel).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

It performs the reference's operation sequence with the reference's dtypes:
float32 grid coordinates (grids.py:63-69), the centred fft2/ifft2 pair with
their scales (utils.py:42-50), exp(i k d) * exp(-i pi lambda d f^2) rebuilt per
step (theory/vacuum.py:7; pathes.py:36-38), the complex64 plane-wave tables and
the (a * E_y) @ E_x product followed by .real (phase_screens.py:125-126, 28),
exp(-1j phi) * u (pathes.py:72-73) and the .item() reductions of measures.py.
"""
import numpy as np


def _axis32(n, delta, column):
    lb = -n // 2 + (n % 2)
    a = np.arange(lb, lb + n, dtype=np.float32) * delta
    return a.reshape(-1, 1) if column else a.reshape(1, -1)


def _fft2(x, delta):
    return np.fft.fftshift(np.fft.fft2(np.fft.fftshift(x))) * delta ** 2


def _ifft2(x, delta):
    n = x.shape[0]
    return np.fft.ifftshift(np.fft.ifft2(np.fft.ifftshift(x))) * (n * delta) ** 2


def _vacuum(u, length, k, delta, n):
    if not length > 0:
        return u
    df = 1 / (n * delta)
    f2 = _axis32(n, df, False) ** 2 + _axis32(n, df, True) ** 2
    out = _ifft2(np.exp(1j * k * length) * np.exp(-1j * np.pi * length * (2 * np.pi / k) * f2) * _fft2(u, delta), df)
    return out.astype(np.complex64)


def _spectrum(rs, base, psd32):
    P = base.shape[0]
    rand = rs.random_sample(size=(1,)).astype(np.float32)
    f_prev = np.insert(base, 0, 0)[:-1]
    rho = np.sqrt(f_prev ** 2 + rand * (base ** 2 - f_prev ** 2))
    theta = 2 * np.pi * rs.random_sample(size=(P,)).astype(np.float32)
    value = (np.array([1, 1j]) @ rs.normal(size=(2, P))).astype(np.complex64) * np.sqrt(psd32)
    return rho, theta, value


def realization(rs, n, delta, wvl, w0, length, positions, base, psd32, radius):
    """One realization: returns (eta in the pupil, mean_x, mean_y, mean_x2, mean_xy, mean_y2)."""
    k = 2 * np.pi / wvl
    x, y = _axis32(n, delta, False), _axis32(n, delta, True)
    u = np.sqrt(2 / np.pi) / w0 * np.exp(-(1 / w0 ** 2 + 1j * 2 * np.pi / wvl / 2 / np.inf) * (x ** 2 + y ** 2))
    for i in range(len(positions)):
        d = positions[i] - positions[i - 1] if i > 0 else positions[0]
        rho, theta, value = _spectrum(rs, base, psd32)
        fx = (rho * np.cos(theta)).reshape((1, -1))
        fy = (rho * np.sin(theta)).reshape((-1, 1))
        phase = (value * np.exp(1j * 2 * np.pi * y @ fy.T) @ np.exp(1j * 2 * np.pi * fx.T @ x)).real
        u = np.exp(-1j * phase) * _vacuum(u, d, k, delta, n)
    u = _vacuum(u, length - positions[-1], k, delta, n)
    I = np.abs(u) ** 2
    d2 = delta ** 2
    mask = x ** 2 + y ** 2 <= radius ** 2
    return ((np.abs(u * mask) ** 2).sum() * d2, (I * x).sum() * d2, (I * (-1) * y).sum() * d2,
            (I * x ** 2).sum() * d2, (I * (-1 * x * y)).sum() * d2, (I * y ** 2).sum() * d2)
