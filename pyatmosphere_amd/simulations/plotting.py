"""Figures for the results: one figure per call, built from plain data
(curves as (x, y, label) triples, transmittance samples as 64-bin densities
over [0, 1], fields as intensity images), returned to the caller and shown
when matplotlib has an interactive backend.  matplotlib is imported lazily:
nothing else in the package needs it."""
from typing import Iterable, Optional, Sequence, Tuple

import numpy as np

Curve = Tuple[Sequence[float], Sequence[float], str]


def _show(fig):
    from matplotlib import pyplot as plt
    if plt.get_backend().lower() not in ("agg", "pdf", "svg", "ps", "cairo", "template"):
        plt.show()
    return fig


def curves(series: Iterable[Curve], xlabel: str, ylabel: str, title: str = "", scatter: bool = False,
           logy: bool = False, ylim: Optional[Tuple[float, float]] = None):
    """One axes with every (x, y, label) series as a line (or markers)."""
    from matplotlib import pyplot as plt
    fig, ax = plt.subplots(figsize=(7, 4.5))
    for x, y, label in series:
        if scatter:
            ax.plot(x, y, "o", ms=3, label=label)
        else:
            ax.plot(x, y, label=label)
    if logy:
        ax.set_yscale("log")
    if ylim is not None:
        ax.set_ylim(*ylim)
    ax.set_xlabel(xlabel)
    ax.set_ylabel(ylabel)
    if title:
        ax.set_title(title)
    ax.grid(alpha=0.3)
    ax.legend(fontsize="small")
    return _show(fig)


def densities(samples: Sequence[Sequence[float]], labels: Sequence[str], bins: int = 64, title: str = ""):
    """Transmittance densities over [0, 1] (`bins` equal bins), all on one axes
    as step outlines so that several apertures can be compared directly."""
    from matplotlib import pyplot as plt
    fig, ax = plt.subplots(figsize=(7, 4.5))
    for s, label in zip(samples, labels):
        s = np.asarray(s, dtype=np.float64)
        h, edges = np.histogram(s, bins=bins, range=(0.0, 1.0), density=True)
        ax.stairs(h, edges, label=f"{label} (n = {len(s)})")
    ax.set_xlim(0.0, 1.0)
    ax.set_xlabel("transmittance $\\eta$")
    ax.set_ylabel("probability density")
    if title:
        ax.set_title(title)
    ax.legend(fontsize="small")
    return _show(fig)


def intensity(image: np.ndarray, extent: Sequence[float], title: str = ""):
    """|u|^2 image over the grid extent (metres), scaled to its maximum."""
    from matplotlib import pyplot as plt
    img = np.asarray(image, dtype=np.float64)
    peak = img.max() if img.size and img.max() > 0 else 1.0
    fig, ax = plt.subplots(figsize=(5.5, 4.5))
    im = ax.imshow(img / peak, extent=extent, origin="upper", cmap="inferno")
    fig.colorbar(im, ax=ax, label="$I / I_{max}$")
    ax.set_xlabel("x, m")
    ax.set_ylabel("y, m")
    if title:
        ax.set_title(title)
    return _show(fig)
