"""Device pipeline of the split-step path: per-channel tables, workspaces and
the batched realization runner that drives libpa_hip.so.

One realization with S screens (pathes.py:61-75) runs as

    row_first -> [col_step(i) -> screen(i) -> row_mid(i)] for i < S -> col_step(S) -> row_last

(from the Gaussian source with S >= 1, source_fresnel -> screen(0) ->
row_mid_source replaces row_first -> col_step(0) -> screen(0) -> row_mid(0):
the source and the transfer function are separable, so the field up to the
first screen is an outer product of two 1-D propagations), where screen(i) = spectral draw (grids.py:98-107; phase_screens.py:93-106),
prep + tables + GEMM (phase_screens.py:108-136).  A batch holds realizations
of one grid size; elements may come from different channels (`chan`) and have
different screen counts (`nsteps`).  Host work is O(N) per channel (1-D
factors in float64) and O(P) per screen when the spectra come from the host
random stream.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr

PSD_GL_NODES = 48


# ---------------------------------------------------------------- host math (O(N), O(P))
def axis_coords(n: int, delta: float) -> np.ndarray:
    """RectGrid axis (grids.py:40-53, 63-69) in float64."""
    lb = -n // 2 + (n % 2)
    return (np.arange(n, dtype=np.float64) + lb) * float(delta)


def log_polar_base(points: int, f_min: float, f_max: float) -> np.ndarray:
    """RandLogPolarGrid.base (grids.py:94-96), float32."""
    return np.exp(np.linspace(np.log(f_min), np.log(f_max), points, dtype=np.float32))


def mvk_psd_n(kappa, Cn2, l0, L0):
    """Modified von Karman index spectrum (theory/models.py:80-85)."""
    k0 = 2 * np.pi / L0
    km = 5.92 / l0
    return 0.033 * Cn2 * np.exp(-(kappa / km) ** 2) / (kappa ** 2 + k0 ** 2) ** (11 / 6)


_GLX, _GLW = np.polynomial.legendre.leggauss(PSD_GL_NODES)


def annulus_psd(base: np.ndarray, k: float, thickness: float, psd_n, panels: int = 4) -> np.ndarray:
    """Phase variance of each log-polar annulus (phase_screens.py:82-91):
    2 pi * integral over [f_{i-1}, f_i] (f_{-1} = 0) of (2 pi)^2 f Phi_phi(f) df with
    Phi_phi(f) = 2 pi k^2 dz Phi_n(2 pi f) (theory/models.py:19-23), by composite
    Gauss-Legendre quadrature in float64.  Returns float64."""
    f = base.astype(np.float64)
    lo = np.concatenate([[0.0], f[:-1]])
    out = np.zeros_like(f)
    edges = np.linspace(0.0, 1.0, panels + 1)
    for p in range(panels):
        a = lo + (f - lo) * edges[p]
        b = lo + (f - lo) * edges[p + 1]
        mid = (0.5 * (a + b))[:, None]
        half = (0.5 * (b - a))[:, None]
        ff = mid + half * _GLX[None, :]
        phi = 2 * np.pi * k ** 2 * thickness * psd_n(2 * np.pi * ff)
        out += half[:, 0] * (((2 * np.pi) ** 2 * ff * phi) @ _GLW)
    return 2 * np.pi * out


def gaussian_factors(n: int, delta: float, wvl: float, w0: float, F0: float) -> Tuple[np.ndarray, np.ndarray]:
    """Separable Gaussian source sqrt(2/pi)/w0 exp(-(1/w0^2 + i k/(2 F0)) r^2)
    (theory/sources.py:16-18) as a row factor gy (carrying the amplitude) and a
    column factor gx, both computed in float64 and rounded to complex64."""
    a = axis_coords(n, delta)
    curv = 0.0 if np.isinf(F0) else 2 * np.pi / wvl / 2 / F0
    g = np.exp(-(1 / w0 ** 2 + 1j * curv) * a ** 2)
    return (np.sqrt(2 / np.pi) / w0 * g).astype(np.complex64), g.astype(np.complex64)


def expi_1d(phase: np.ndarray) -> np.ndarray:
    """exp(-i phase) in float64, rounded to complex64."""
    return np.exp(-1j * phase).astype(np.complex64)


@dataclass
class ChannelSpec:
    """Everything the device path needs to know about one channel."""
    n: int
    delta: float
    wvl: float
    w0: float
    F0: float
    length: float
    positions: np.ndarray            # screen positions (pathes.py:80-89)
    points: int                      # spectral points per screen
    base: np.ndarray                 # float32 (P,)
    psd: np.ndarray                  # float32 (P,), stored as the reference stores it
    pupils: List[Tuple[float, float, float]] = field(default_factory=list)  # (radius, sx, sy)
    shift: Tuple[float, float] = (0.0, 0.0)  # screen coordinate shift (phase_screens.py:112-113)
    gains: Optional[List[float]] = None      # amplitude factor applied with each Fresnel step

    @property
    def k(self) -> float:
        return 2 * np.pi / self.wvl

    @property
    def nscreens(self) -> int:
        return len(self.positions)

    def step_lengths(self) -> List[float]:
        """Fresnel steps d_0 .. d_S (pathes.py:67-75)."""
        pos = list(self.positions)
        d = [pos[i] - pos[i - 1] if i > 0 else pos[0] for i in range(len(pos))]
        d.append(self.length - (pos[-1] if pos else 0.0))
        return [float(v) for v in d]

    def source_factors(self) -> Tuple[np.ndarray, np.ndarray]:
        return gaussian_factors(self.n, self.delta, self.wvl, self.w0, self.F0)

    def transfer_factors(self, steps: int) -> np.ndarray:
        """(steps, 2, N) complex64: row factor h(r), column factor s h(c) with
        h(m) = exp(-i pi lambda d f_m^2) on the unshifted frequency axis and
        s = exp(i k d)/N^2 (theory/vacuum.py:7 plus the inverse-FFT norm).
        d <= 0 steps are the identity (pathes.py:30, 39-40)."""
        n = self.n
        f = np.fft.fftfreq(n, d=self.delta)
        out = np.zeros((steps, 2, n), np.complex64)
        for i, d in enumerate(self.step_lengths()):
            if d > 0:
                h = np.exp(-1j * np.pi * d * self.wvl * f ** 2)
                s = np.exp(1j * self.k * d) / n ** 2
            else:
                h = np.ones(n)
                s = 1.0 / n ** 2
            if self.gains is not None:
                s = s * self.gains[i]
            out[i, 0] = h.astype(np.complex64)
            out[i, 1] = (s * h).astype(np.complex64)
        for i in range(len(self.step_lengths()), steps):
            out[i, 0] = 1.0
            out[i, 1] = 1.0 / n ** 2
        return out


# ---------------------------------------------------------------- device tables
_TWIDDLES: Dict[Tuple[int, int], torch.Tensor] = {}


def twiddles(n: int, device: torch.device) -> torch.Tensor:
    key = (n, device.index or 0)
    t = _TWIDDLES.get(key)
    if t is None:
        t = torch.empty(n, dtype=torch.complex64, device=device)
        call("pa_fill_twiddles", n, ptr(t), _lib.stream_handle())
        _TWIDDLES[key] = t
    return t


class DeviceTables:
    """Per-channel tables stacked for the channels of one batch (same N, P)."""

    def __init__(self, specs: Sequence[ChannelSpec], device: torch.device):
        assert specs, "no channels"
        n, P = specs[0].n, specs[0].points
        for s in specs:
            if s.n != n or s.points != P:
                raise ValueError("channels of one batch must share grid size and spectral points")
            if len(s.pupils) > _lib.PA_MAX_PUPILS:
                raise ValueError(f"at most {_lib.PA_MAX_PUPILS} pupils per channel")
        self.specs = list(specs)
        self.n, self.points, self.device = n, P, device
        self.max_screens = max(s.nscreens for s in specs)
        self.max_steps = self.max_screens + 1
        C = len(specs)
        src = np.zeros((C, 2, n), np.complex64)
        htab = np.zeros((C, self.max_steps, 2, n), np.complex64)
        chp = np.zeros((C, _lib.PA_CH_PARAMS), np.float64)
        pup = np.zeros((C, _lib.PA_MAX_PUPILS, 3), np.float64)
        base = np.zeros((C, P), np.float32)
        sq = np.zeros((C, P), np.float32)
        for i, s in enumerate(specs):
            src[i, 0], src[i, 1] = s.source_factors()
            htab[i] = s.transfer_factors(self.max_steps)
            chp[i, :3] = (s.delta, s.shift[0], s.shift[1])
            for j, p in enumerate(s.pupils):
                pup[i, j] = p
            base[i] = s.base
            sq[i] = np.sqrt(s.psd.astype(np.float32))
        self.npup = max(len(s.pupils) for s in specs)
        to = lambda a: torch.from_numpy(a).to(device)
        self.src, self.htab, self.chparam, self.pup = to(src), to(htab), to(chp), to(pup)
        self.base, self.sqrt_psd = to(base), to(sq)


class Workspace:
    """Device buffers for `batch` realizations of an n x n grid with P points."""

    def __init__(self, n: int, points: int, batch: int, mode: str, device: torch.device, screens: int = 1):
        self.n, self.points, self.batch, self.device = n, points, batch, device
        self.mode_name = mode
        self.mode = _lib.GEMM_MODES[mode]
        # launch shapes of the screen kernels: the autotuner's cached choice for
        # this (n, points, batch, mode) on this device, else 0 (library default)
        from . import autotune
        shape = autotune.lookup(n, points, batch, mode, device)
        self.gemm_tile = shape["gemm_tile"] if shape else 0
        self.table_rows = shape["table_rows"] if shape else 0
        self.col_group = shape.get("col_group", 0) if shape else 0
        T = call("pa_tile_rows", n)
        self.tiles = n // T
        opb = call("pa_operand_bytes", n, points, self.mode)
        B = batch
        e = lambda *shape, dt: torch.empty(*shape, dtype=dt, device=device)
        self.work = e(B, n, n, dt=torch.complex64)
        # float32 screen angle of the current screen (work layout): phi2 plus the
        # row and column angles; the row pass applies e^{-i ang}
        self.ang = e(B, n, n, dt=torch.float32)
        self.phi2: Optional[torch.Tensor] = None   # real phi2, allocated when a screen is composed
        # per-screen angles and row/column factors for the overlapped schedule (grown on demand)
        self.ang_s: List[torch.Tensor] = []
        self.esc_s: List[torch.Tensor] = []
        self.side = None
        self.opa = e(B, opb, dt=torch.uint8)
        self.opb = e(B, opb, dt=torch.uint8)
        self.fxy = e(B, points, 4, dt=torch.float64)   # fx, fy, amplitude with the screen shift folded in
        self.sparam = e(B, _lib.PA_SCREEN_PARAMS, dt=torch.float64)
        self.hg = e(B, 2, n, dt=torch.float64)
        self.esc = e(B, 2, n, dt=torch.complex64)
        self.sep = e(B, 2, n, dt=torch.complex64)     # separable first step: uy, ux per element
        self.rho = e(B, points, dt=torch.float32)
        self.theta = e(B, points, dt=torch.float32)
        self.value = e(B, points, dt=torch.complex64)
        self.part = e(B, self.tiles, _lib.PA_PART_SUMS, dt=torch.float64)
        self.moments = e(B, _lib.PA_NUM_MOMENTS, dt=torch.float64)
        self.rid = e(B, dt=torch.int64)
        self.chan = e(B, dt=torch.int32)
        self.nsteps = e(B, dt=torch.int32)

    def phi2_buffer(self) -> torch.Tensor:
        if self.phi2 is None:
            self.phi2 = torch.empty(self.batch, self.n, self.n, dtype=torch.float32, device=self.device)
        return self.phi2

    def screen_buffers(self, screens: int) -> None:
        while len(self.ang_s) < screens:
            self.ang_s.append(torch.empty(self.batch, self.n, self.n, dtype=torch.float32, device=self.device))
            self.esc_s.append(torch.empty(self.batch, 2, self.n, dtype=torch.complex64, device=self.device))
        if self.side is None:
            self.side = torch.cuda.Stream(device=self.device)   # default priority

    def synth_buffers(self, screens: int) -> "SynthBuffers":
        """Buffers for synthesizing `screens` screens of a whole batch at once
        (SynthBuffers), allocated on first use and kept."""
        sb = getattr(self, "_synth", None)
        if sb is None or sb.screens < screens:
            self._synth = None
            sb = self._synth = SynthBuffers(self, screens)
        return sb

    def bytes(self) -> int:
        return sum(t.numel() * t.element_size() for t in vars(self).values() if isinstance(t, torch.Tensor))


# Whole-batch screen synthesis (SynthBuffers) is used while its buffers stay
# below this many bytes; larger batches synthesize screen by screen.
SYNTH_BATCH_MAX_BYTES = 16 << 30


def synth_bytes_per_element(n: int, points: int, mode: int) -> int:
    """Device bytes of one (screen, realization) pair in SynthBuffers."""
    opb = call("pa_operand_bytes", n, points, mode)
    return 2 * opb + points * 32 + _lib.PA_SCREEN_PARAMS * 8 + 2 * n * (8 + 8) + n * n * 4 + 12


class SynthBuffers:
    """Screen synthesis of every screen of a batch as one virtual batch:
    element v = i * B + b is screen i of realization b (B the run's batch).
    The device draw runs per screen into its slice (it takes the screen index);
    the plane-wave tables and the GEMM then run once over all S * B pairs.
    Every pair is computed exactly as in the screen-by-screen schedule, so the
    angles are bit-identical; what changes is five times fewer launches of the
    small latency-bound kernels, each with a five times larger grid.  Launch
    shapes are the library's defaults for the virtual batch."""

    def __init__(self, ws: "Workspace", screens: int):
        n, P, V = ws.n, ws.points, screens * ws.batch
        self.n, self.points, self.mode, self.mode_name = n, P, ws.mode, ws.mode_name
        self.table_rows, self.gemm_tile = 0, 0
        self.screens, self.capacity = screens, V
        opb = call("pa_operand_bytes", n, P, ws.mode)
        e = lambda *shape, dt: torch.empty(*shape, dtype=dt, device=ws.device)
        self.opa = e(V, opb, dt=torch.uint8)
        self.opb = e(V, opb, dt=torch.uint8)
        self.fxy = e(V, P, 4, dt=torch.float64)
        self.sparam = e(V, _lib.PA_SCREEN_PARAMS, dt=torch.float64)
        self.hg = e(V, 2, n, dt=torch.float64)
        self.esc = e(V, 2, n, dt=torch.complex64)
        self.ang = e(V, n, n, dt=torch.float32)
        self.chan = e(V, dt=torch.int32)
        self.active = e(V, dt=torch.int32)

    def draw_view(self, ws: "Workspace", screen: int, batch: int):
        """The workspace as the device draw of screen `screen` sees it: the
        batch's channels and realization ids, this screen's fxy and sparam slice."""
        lo, hi = screen * batch, (screen + 1) * batch
        return _View(points=self.points, mode=self.mode, chan=ws.chan, rid=ws.rid,
                     fxy=self.fxy[lo:hi], sparam=self.sparam[lo:hi])


class _View:
    def __init__(self, **kw):
        self.__dict__.update(kw)


# ---------------------------------------------------------------- pipeline pieces
def prep_screen(ws: Workspace, tables: DeviceTables, step: int, batch: int, nsteps: Optional[torch.Tensor],
                stream) -> None:
    """Frequencies, shift-rotated amplitudes and GEMM operand scales of screen
    `step` from ws.rho/theta/value and the elements' channels."""
    call("pa_screen_prep", ws.points, ptr(ws.rho), ptr(ws.theta), ptr(ws.value), ptr(ws.fxy), ptr(ws.sparam),
         ws.mode, ptr(tables.chparam), ptr(ws.chan), ptr(nsteps), step, batch, stream)


def plane_wave_tables(ws: Workspace, tables: DeviceTables, step: int, batch: int, nsteps: Optional[torch.Tensor],
                      stream, esc: Optional[torch.Tensor] = None, rows: Optional[int] = None) -> None:
    """GEMM operands, float64 1-D parts h, g and the row/column factors of screen `step`."""
    call("pa_screen_tables", ws.n, ws.points, ptr(ws.fxy), ptr(ws.sparam), ptr(tables.chparam),
         ptr(ws.chan), ptr(ws.opa), ptr(ws.opb), ptr(ws.hg), ptr(ws.esc if esc is None else esc), ws.mode,
         ws.table_rows if rows is None else rows, ptr(nsteps), step, batch, stream)


def screen_gemm(ws: Workspace, step: int, batch: int, nsteps: Optional[torch.Tensor], stream,
                fac: Optional[torch.Tensor] = None, esc: Optional[torch.Tensor] = None,
                phi2: Optional[torch.Tensor] = None, ang: Optional[torch.Tensor] = None,
                tile: Optional[int] = None) -> None:
    """phi2 = A B^T of screen `step`, into any of: the screen angle `ang` (what
    the row pass applies), the real phi2 `phi2`, the complex factor `fac`."""
    call("pa_screen_gemm", ws.n, ws.points, ptr(ws.opa), ptr(ws.opb), ptr(ws.sparam),
         ptr(ws.esc if esc is None else esc), ptr(phi2), ptr(ang), ptr(fac), ws.mode,
         ws.gemm_tile if tile is None else tile, ptr(nsteps), step, batch, stream)


def synth_screen(ws: Workspace, tables: DeviceTables, step: int, batch: int,
                 nsteps: Optional[torch.Tensor], stream, fac: Optional[torch.Tensor] = None,
                 esc: Optional[torch.Tensor] = None, phi2: Optional[torch.Tensor] = None,
                 ang: Optional[torch.Tensor] = None, prepped: bool = False) -> None:
    """Screen `step` of every active element from ws.rho/theta/value
    (phase_screens.py:108-136): prep, plane-wave tables, GEMM.  The GEMM writes
    the screen angle into `ang` (default ws.ang), and the real phi2 and the
    complex factor eh eg e^{-i phi2} into `phi2` / `fac` when given."""
    ang = ws.ang if ang is None else ang
    if not prepped:   # (else draw_prep_device has done it)
        prep_screen(ws, tables, step, batch, nsteps, stream)
    plane_wave_tables(ws, tables, step, batch, nsteps, stream, esc=esc)
    screen_gemm(ws, step, batch, nsteps, stream, fac=fac, esc=esc, phi2=phi2, ang=ang)


def draw_prep_device(ws: Workspace, tables: DeviceTables, seed: int, screen: int, batch: int,
                     nsteps: Optional[torch.Tensor], stream) -> None:
    """draw_device followed by prep_screen in one launch (the same fxy and
    sparam; ws.rho/theta/value are not written)."""
    call("pa_screen_draw_prep", ws.points, ptr(tables.base), ptr(tables.sqrt_psd), ptr(ws.chan), ptr(ws.rid),
         int(seed) & (2 ** 64 - 1), screen, ptr(ws.fxy), ptr(ws.sparam), ws.mode, ptr(tables.chparam), ptr(nsteps),
         screen, batch, stream)


def draw_device(ws: Workspace, tables: DeviceTables, seed: int, screen: int, batch: int, stream) -> None:
    """Device Philox spectral draw of screen `screen` for realizations ws.rid."""
    call("pa_spectral_draw", ws.points, ptr(tables.base), ptr(tables.sqrt_psd), ptr(ws.chan), ptr(ws.rid),
         int(seed) & (2 ** 64 - 1), screen, ptr(ws.rho), ptr(ws.theta), ptr(ws.value), batch, stream)


def upload(dst: torch.Tensor, arr, dtype) -> None:
    """A host array into a device tensor without blocking the host: the data
    are copied into pinned staging memory now (PyTorch's caching host allocator
    keeps it until the transfer has run) and the transfer is ordered on the
    current stream ahead of the kernels launched after it.  A pageable copy
    would make the host wait for all earlier work on the stream, leaving the
    device idle while the next launches are issued."""
    dst.copy_(torch.from_numpy(np.ascontiguousarray(arr, dtype)).pin_memory(), non_blocking=True)


def upload_spectra(ws: Workspace, rho: np.ndarray, theta: np.ndarray, value: np.ndarray, batch: int) -> None:
    """Host-drawn spectra (B, P) of one screen into the workspace."""
    upload(ws.rho[:batch], rho, np.float32)
    upload(ws.theta[:batch], theta, np.float32)
    upload(ws.value[:batch], value, np.complex64)


class StepRunner:
    """Step-level driver of one batch (used by the generator API, which yields
    the field after every screen, and by run_batch)."""

    def __init__(self, ws: Workspace, tables: DeviceTables, chan: np.ndarray, nscreens: np.ndarray,
                 rid: Optional[np.ndarray] = None):
        self.ws, self.tables = ws, tables
        self.B = B = len(chan)
        if B > ws.batch:
            raise ValueError("batch larger than the workspace")
        self.stream = _lib.stream_handle()
        self.tw = twiddles(ws.n, ws.device)
        upload(ws.chan[:B], chan, np.int32)
        upload(ws.nsteps[:B], nscreens, np.int32)
        uniform = bool(np.all(np.asarray(nscreens) == tables.max_screens))
        self.ns = None if uniform else ws.nsteps
        if rid is not None:
            upload(ws.rid[:B], rid, np.int64)
        self.smax = int(np.max(nscreens)) if len(nscreens) else 0
        self.smin = int(np.min(nscreens)) if len(nscreens) else 0
        self.sep0 = False
        self.chan_np, self.nscreens_np = np.asarray(chan), np.asarray(nscreens)

    def whole_batch_synthesis(self) -> bool:
        """Whether synthesize_all fits SYNTH_BATCH_MAX_BYTES for this batch."""
        ws = self.ws
        need = self.smax * self.B * synth_bytes_per_element(ws.n, ws.points, ws.mode)
        return self.smax > 0 and need <= SYNTH_BATCH_MAX_BYTES

    def synthesize_all(self, seed: int, stream=None, timed=None, upload_ids: bool = True) -> List[torch.Tensor]:
        """Every screen's synthesis for the batch as one virtual batch
        (SynthBuffers): the device draw per screen, then the tables and the GEMM
        once over all (screen, realization) pairs.  Returns the screen angles
        of screen i as a (B, n, n) view.  `timed(name, fn)`, when given, runs
        each launch (a timing hook).  Without `upload_ids` the pairs' channels
        and activity flags are taken as an earlier call left them (graph
        capture records no host uploads)."""
        ws, B, S = self.ws, self.B, self.smax
        sb = ws.synth_buffers(S)
        st = self.stream if stream is None else stream
        run = timed if timed is not None else (lambda name, fn: fn())
        if upload_ids:
            upload(sb.chan[:S * B], np.tile(np.asarray(self.chan_np, np.int32), S), np.int32)
            if self.ns is not None:
                act = (np.arange(S)[:, None] < np.asarray(self.nscreens_np)[None, :]).astype(np.int32).reshape(-1)
                upload(sb.active[:S * B], act, np.int32)
        nv = None if self.ns is None else sb.active
        for i in range(S):
            run("screen_draw_prep",
                lambda i=i: draw_prep_device(sb.draw_view(ws, i, B), self.tables, seed, i, B, self.ns, st))
        run("screen_tables", lambda: plane_wave_tables(sb, self.tables, 0, S * B, nv, st))
        run("screen_gemm", lambda: screen_gemm(sb, 0, S * B, nv, st, ang=sb.ang))
        return [sb.ang[i * B:(i + 1) * B] for i in range(S)]

    def begin(self, field_in: Optional[torch.Tensor] = None, separable: bool = False) -> None:
        """The source (or `field_in`) into the work buffer.  With `separable`,
        no input field and a first screen in every element, the Gaussian
        source's first Fresnel step is taken as two 1-D propagations instead
        (pa_source_fresnel): fresnel(0) then has nothing to do and apply(0)
        forms the field of the first screen plane from the two vectors."""
        ws, t = self.ws, self.tables
        self.sep0 = bool(separable and field_in is None and self.smin >= 1)
        if self.sep0:
            call("pa_source_fresnel", ws.n, ptr(ws.sep), ptr(t.src), ptr(t.htab), ptr(ws.chan), t.max_steps,
                 ptr(self.tw), self.B, self.stream)
            return
        call("pa_row_first", ws.n, ptr(ws.work), ptr(field_in), None if field_in is not None else ptr(t.src),
             ptr(ws.chan), ptr(self.tw), self.B, self.stream)

    def fresnel(self, step: int) -> None:
        ws, t = self.ws, self.tables
        if step == 0 and self.sep0:
            return          # taken by begin(separable=True)
        call("pa_col_step", ws.n, ptr(ws.work), ptr(t.htab), ptr(ws.chan), ptr(self.ns), step, t.max_steps,
             ptr(self.tw), ws.col_group, self.B, self.stream)

    def screen(self, step: int, *, seed: Optional[int] = None,
               host: Optional[Tuple[np.ndarray, np.ndarray, np.ndarray]] = None,
               compose: bool = False) -> Optional[torch.Tensor]:
        phi = self.synthesize(step, seed=seed, host=host, compose=compose)
        self.apply(step)
        return phi

    def synthesize(self, step: int, *, seed: Optional[int] = None,
                   host: Optional[Tuple[np.ndarray, np.ndarray, np.ndarray]] = None, compose: bool = False,
                   ang: Optional[torch.Tensor] = None, esc: Optional[torch.Tensor] = None,
                   stream=None) -> Optional[torch.Tensor]:
        """Spectra, tables and GEMM of screen `step` into (ang, esc) (default:
        the workspace's single buffers) on `stream` (default: the runner's).  With
        `compose`, also returns the full phase screen (B, n, n) float32."""
        ws = self.ws
        st = self.stream if stream is None else stream
        if host is not None:
            upload_spectra(ws, *host, batch=self.B)
        else:
            draw_prep_device(ws, self.tables, seed, step, self.B, self.ns, st)
        p2 = ws.phi2_buffer() if compose else None
        synth_screen(ws, self.tables, step, self.B, self.ns, st, esc=esc, phi2=p2, ang=ang, prepped=host is None)
        phi = None
        if compose:
            phi = torch.empty(self.B, ws.n, ws.n, dtype=torch.float32, device=ws.device)
            call("pa_screen_compose", ws.n, ptr(ws.sparam), ptr(ws.hg), ptr(p2), ptr(phi), self.B, st)
        return phi

    def apply(self, step: int, ang: Optional[torch.Tensor] = None) -> None:
        """Inverse row FFT, screen multiply, forward row FFT of screen `step`."""
        ws = self.ws
        a = ws.ang if ang is None else ang
        if step == 0 and self.sep0:
            call("pa_row_mid_source", ws.n, ptr(ws.work), ptr(ws.sep), ptr(a), ptr(self.tw), self.B, self.stream)
            return
        call("pa_row_mid", ws.n, ptr(ws.work), ptr(a), ptr(self.ns), step, ptr(self.tw), self.B, self.stream)

    def finish(self, field_out: Optional[torch.Tensor] = None, reduce: bool = True) -> Optional[torch.Tensor]:
        ws, t = self.ws, self.tables
        part = ws.part if reduce else None
        call("pa_row_last", ws.n, ptr(ws.work), ptr(field_out), ptr(part), ptr(t.chparam), ptr(t.pup), t.npup,
             None, ptr(ws.chan), ptr(self.tw), self.B, self.stream)
        if reduce:
            call("pa_moments_finalize", ws.n, ptr(ws.part), ptr(t.chparam), t.npup, ptr(ws.chan),
                 ptr(ws.moments), self.B, self.stream)
            return ws.moments[:self.B]
        return None

    def tracked(self, radii: Sequence[float]) -> torch.Tensor:
        """Transmittance through pupils of the given radii centred on each
        element's own beam centroid (TrackedPDTResult, simulations/pdt.py:51-66):
        a second inverse row pass over the work buffer with the centroids of
        finish()'s moments, so no field is stored.  Returns (B, PA_NUM_MOMENTS)
        moments whose [7 + p] entries are the tracked transmittances."""
        ws, t = self.ws, self.tables
        if len(radii) > _lib.PA_MAX_PUPILS:
            raise ValueError(f"at most {_lib.PA_MAX_PUPILS} tracked pupils")
        pup = np.zeros((1, _lib.PA_MAX_PUPILS, 3))
        pup[0, :len(radii), 0] = radii
        pup_t = torch.from_numpy(np.repeat(pup, len(t.specs), axis=0)).to(ws.device)
        centroid = ws.moments[:self.B].clone()
        out = torch.empty_like(centroid)
        call("pa_row_last", ws.n, ptr(ws.work), None, ptr(ws.part), ptr(t.chparam), ptr(pup_t), len(radii),
             ptr(centroid), ptr(ws.chan), ptr(self.tw), self.B, self.stream)
        call("pa_moments_finalize", ws.n, ptr(ws.part), ptr(t.chparam), len(radii), ptr(ws.chan), ptr(out),
             self.B, self.stream)
        return out


def run_batch(ws: Workspace, tables: DeviceTables, chan: np.ndarray, nscreens: np.ndarray, *,
              seed: Optional[int] = None, rid: Optional[np.ndarray] = None,
              host_spectra: Optional[Sequence[Tuple[np.ndarray, np.ndarray, np.ndarray]]] = None,
              field_in: Optional[torch.Tensor] = None, field_out: Optional[torch.Tensor] = None,
              reduce: bool = True, screens_out: Optional[List[torch.Tensor]] = None,
              overlap: Optional[bool] = None, tracked: Optional[Sequence[float]] = None,
              separable: Optional[bool] = None):
    """Propagate `len(chan)` realizations through their channels.

    Spectra come from the device Philox stream (`seed`, `rid`) or from
    `host_spectra[i] = (rho, theta, value)` arrays of shape (B, P) per screen i.
    Returns ws.moments[:B] (device) when `reduce`, which holds per element
    eta, mean_x, mean_y, mean_x2, mean_xy, mean_y2, mean_x2_r, eta per pupil.
    With `tracked` (pupil radii), returns (moments, tracked moments) where the
    latter's [7 + p] entries are the transmittances through pupils centred on
    each element's centroid (StepRunner.tracked).  `separable` (default
    SEPARABLE_SOURCE) takes the source's first Fresnel step as two 1-D
    propagations when there is no input field and every element has a screen
    (StepRunner.begin).
    """
    run = StepRunner(ws, tables, chan, nscreens, rid)
    sep = SEPARABLE_SOURCE if separable is None else bool(separable)
    if overlap is None:
        overlap = use_overlap(ws, len(chan), int(np.max(nscreens)) if len(nscreens) else 0)
    device_draws = host_spectra is None and screens_out is None and run.smax > 0
    whole = device_draws and SYNTH_WHOLE_BATCH and run.whole_batch_synthesis()
    if device_draws and overlap:
        overlapped_body(run, seed, field_in, sep, whole)
    elif whole:
        angs = run.synthesize_all(seed)
        run.begin(field_in, sep)
        for i in range(run.smax + 1):
            run.fresnel(i)
            if i < run.smax:
                run.apply(i, ang=angs[i])
    else:
        run.begin(field_in, sep)
        for i in range(run.smax + 1):
            run.fresnel(i)
            if i < run.smax:
                phi = run.screen(i, seed=seed, host=None if host_spectra is None else host_spectra[i],
                                 compose=screens_out is not None)
                if screens_out is not None:
                    screens_out.append(phi)
    mom = run.finish(field_out, reduce or bool(tracked))
    if tracked:
        return mom, run.tracked(tracked)
    return mom


OVERLAP_DEFAULT = True
# device-drawn batches synthesize all their screens at once (SynthBuffers)
# when the buffers fit SYNTH_BATCH_MAX_BYTES
SYNTH_WHOLE_BATCH = True
# the Gaussian source's first Fresnel step as two 1-D propagations (exact:
# source and transfer function are separable; StepRunner.begin)
SEPARABLE_SOURCE = True


def _overlap_rule(batch: int, n: int) -> bool:
    """Default: the side stream pays off when one batch does not fill the GPU."""
    return batch * n * n <= (1 << 23)


OVERLAP_RULE = _overlap_rule


def use_overlap(ws: "Workspace", batch: int, screens: int) -> bool:
    """Whether screen synthesis of a batch runs on the side stream: the
    autotuner's cached choice for this batch size, else the default rule."""
    if not OVERLAP_DEFAULT:
        return False
    if OVERLAP_RULE is _overlap_rule:
        from . import autotune
        e = autotune.lookup_batch(ws.n, ws.points, screens, ws.mode_name, ws.device)
        if e is not None and e["batch"] == batch:
            return e["overlap"]
    return OVERLAP_RULE(batch, ws.n)


def overlapped_body(run: "StepRunner", seed: int, field_in: Optional[torch.Tensor], separable: bool = False,
                    whole: bool = False, upload_ids: bool = True) -> None:
    """Two-stream schedule of one batch: every screen's synthesis (spectral draw,
    tables, GEMM: MFMA-bound) runs on a side stream into its own buffers while
    the Fresnel passes (latency/HBM-bound) run on the main stream; row_mid(i)
    waits on an event recorded after screen i.  With `whole`, the side stream
    synthesizes all screens at once (StepRunner.synthesize_all) and the first
    row_mid waits for it.  Streams keep default priority."""
    ws = run.ws
    # buffers are allocated here, on the main stream, not under the side stream
    if whole:
        ws.synth_buffers(run.smax)
        ws.screen_buffers(0)
    else:
        ws.screen_buffers(run.smax)
    main = torch.cuda.current_stream()
    side = ws.side
    side.wait_stream(main)
    events, angs = [], []
    with torch.cuda.stream(side):
        st = _lib.stream_handle()
        if whole:
            angs = run.synthesize_all(seed, stream=st, upload_ids=upload_ids)
            ev = torch.cuda.Event()
            ev.record(side)
            events = [ev] * run.smax
        else:
            for i in range(run.smax):
                run.synthesize(i, seed=seed, ang=ws.ang_s[i], esc=ws.esc_s[i], stream=st)
                ev = torch.cuda.Event()
                ev.record(side)
                events.append(ev)
            angs = ws.ang_s
    run.begin(field_in, separable)
    for i in range(run.smax + 1):
        run.fresnel(i)
        if i < run.smax:
            main.wait_event(events[i])
            run.apply(i, ang=angs[i])
    main.wait_stream(side)


class GraphedBatch:
    """A batch configuration (workspace, tables, channel per element, screen
    counts, seed) whose whole launch sequence is captured once into a HIP graph
    and replayed for each new set of realization ids.  Only the device-RNG path
    is captured (host spectra need host uploads); the realization ids are copied
    into the workspace before each replay, outside the graph, and the graph
    holds no memset nodes (the kernels overwrite everything they read)."""

    def __init__(self, ws: Workspace, tables: DeviceTables, chan: np.ndarray, nscreens: np.ndarray, seed: int,
                 field_out: Optional[torch.Tensor] = None):
        self.ws, self.tables, self.seed, self.field_out = ws, tables, int(seed), field_out
        self.B = len(chan)
        self.chan = np.asarray(chan)
        self.nscreens = np.asarray(nscreens)
        self.graph = None
        twiddles(ws.n, ws.device)

    def run(self, rid: np.ndarray) -> torch.Tensor:
        rid = np.asarray(rid, np.int64)
        if len(rid) != self.B:
            raise ValueError("rid length must equal the captured batch")
        if self.graph is None:
            out = run_batch(self.ws, self.tables, self.chan, self.nscreens, seed=self.seed, rid=rid,
                            field_out=self.field_out)      # eager warm-up, also uploads chan/nsteps
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._capture()
            self.graph = g
            torch.cuda.synchronize()
            return out
        upload(self.ws.rid[:self.B], rid, np.int64)
        self.graph.replay()
        return self.ws.moments[:self.B]

    def _capture(self, overlap: Optional[bool] = None):
        """The launch sequence of run_batch with the per-batch uploads left out."""
        ws, t = self.ws, self.tables
        run = StepRunner.__new__(StepRunner)
        run.ws, run.tables, run.B = ws, t, self.B
        run.stream = _lib.stream_handle()
        run.tw = twiddles(ws.n, ws.device)
        uniform = bool(np.all(self.nscreens == t.max_screens))
        run.ns = None if uniform else ws.nsteps
        run.smax = int(np.max(self.nscreens))
        run.smin = int(np.min(self.nscreens))
        run.sep0 = False
        run.chan_np, run.nscreens_np = self.chan, self.nscreens
        if overlap is None:
            overlap = use_overlap(ws, self.B, run.smax)
        # the eager warm-up (run_batch) took the same decision and left the
        # pairs' channels and activity flags in the synthesis buffers
        whole = run.smax > 0 and SYNTH_WHOLE_BATCH and run.whole_batch_synthesis()
        if run.smax > 0 and overlap:
            overlapped_body(run, self.seed, None, SEPARABLE_SOURCE, whole, upload_ids=False)
        elif whole:
            angs = run.synthesize_all(self.seed, upload_ids=False)
            run.begin(None, SEPARABLE_SOURCE)
            for i in range(run.smax + 1):
                run.fresnel(i)
                if i < run.smax:
                    run.apply(i, ang=angs[i])
        else:
            run.begin(None, SEPARABLE_SOURCE)
            for i in range(run.smax + 1):
                run.fresnel(i)
                if i < run.smax:
                    run.screen(i, seed=self.seed)
        run.finish(self.field_out, True)


class Ensemble:
    """Per-channel Monte Carlo ensembles kept on the device: float64 sums (and
    sums of squares) of the BeamResult quantities and 64-bin integer
    transmittance histograms per pupil (simulations/beam.py:35-71;
    simulations/pdt.py:28-48), fed one batch of finalized moments at a time by
    pa_ensemble_accumulate with no host round trip.  A channel stops taking
    beam samples at `beam_samples` and PDT samples at `pdt_samples`."""

    QUANTITIES = ("eta", "mean_x", "mean_y", "mean_x2", "mean_xy", "mean_y2", "mean_x2_r",
                  "bw2_x", "bw2_y", "lt2", "st2")

    def __init__(self, channels: int, npup: int, beam_samples: int, pdt_samples: int, device: torch.device):
        if npup > _lib.PA_MAX_PUPILS:
            raise ValueError(f"at most {_lib.PA_MAX_PUPILS} pupils")
        self.channels, self.npup = int(channels), int(npup)
        self.beam_samples, self.pdt_samples = int(beam_samples), int(pdt_samples)
        self.acc = torch.zeros(channels, 2, _lib.PA_ENS_QUANTITIES, dtype=torch.float64, device=device)
        self.counts = torch.zeros(channels, 2, dtype=torch.int64, device=device)
        # uint64 counts, held as int64 (never near 2^63)
        self.hist = torch.zeros(channels, _lib.PA_MAX_PUPILS, _lib.PA_PDT_BINS, dtype=torch.int64, device=device)

    def add(self, moments: torch.Tensor, slot_channel: Optional[torch.Tensor]) -> None:
        """Accumulate a batch of finalized moments (B, PA_NUM_MOMENTS); element b
        belongs to channel slot_channel[b] (int32 device tensor; None = 0)."""
        call("pa_ensemble_accumulate", moments.shape[0], ptr(moments), ptr(slot_channel), self.channels, self.npup,
             self.beam_samples, self.pdt_samples, ptr(self.acc), ptr(self.counts), ptr(self.hist),
             _lib.stream_handle())

    def complete(self) -> bool:
        c = self.counts.cpu().numpy()
        return bool(np.all(c[:, 0] >= self.beam_samples) and np.all(c[:, 1] >= self.pdt_samples))

    def beam(self) -> List[Dict[str, Tuple[float, float]]]:
        """Per channel: bw, lt, st as BeamResult reports them (sqrt of the mean
        and its error std(ddof=1)/sqrt(n)/2/sqrt(mean), beam.py:49-71), plus the
        sample count."""
        acc = self.acc.cpu().numpy()
        n = self.counts.cpu().numpy()[:, 0].astype(np.float64)
        out = []
        q = {k: i for i, k in enumerate(self.QUANTITIES)}
        for c in range(self.channels):
            row = {"count": int(n[c])}
            for name, key in (("bw", "bw2_x"), ("lt", "lt2"), ("st", "st2")):
                s1, s2 = acc[c, 0, q[key]], acc[c, 1, q[key]]
                mean = s1 / n[c]
                var = (s2 - s1 * s1 / n[c]) / (n[c] - 1) if n[c] > 1 else np.nan
                m = np.sqrt(mean)
                row[name] = (float(m), float(np.sqrt(max(var, 0.0)) / np.sqrt(n[c]) / 2 / m))
            out.append(row)
        return out

    def histogram(self) -> np.ndarray:
        """(channels, npup, 64) integer counts over [0, 1]."""
        return self.hist.cpu().numpy()[:, :self.npup]

    def pdt(self) -> Tuple[np.ndarray, np.ndarray]:
        """Densities (channels, npup, 64) normalised as numpy.histogram(density=True)
        and the 65 bin edges."""
        h = self.histogram().astype(np.float64)
        edges = np.linspace(0.0, 1.0, _lib.PA_PDT_BINS + 1)
        tot = h.sum(axis=-1, keepdims=True)
        with np.errstate(invalid="ignore", divide="ignore"):
            return h / tot / np.diff(edges), edges


class MixedBatch:
    """Slots of mixed channels (grid size, spacing, screen count, turbulence,
    wavelength, beam) advanced together by one captured HIP graph.

    Slots are grouped into shape classes by (grid size, spectral points); each
    class has its own workspace and channel tables, and its launch sequence is
    captured on its own stream, forked from and joined to the capturing stream,
    so the classes' kernels run concurrently within one graph replay.  One
    `run(rid)` advances every slot by one realization (device Philox draws
    keyed by rid[slot]) and returns the moments (slots, PA_NUM_MOMENTS) in slot
    order; per slot they equal a run_batch of that slot's class."""

    def __init__(self, specs: Sequence[ChannelSpec], slot_channel: Sequence[int], seed: int, mode: str,
                 device: torch.device):
        self.specs, self.seed, self.device = list(specs), int(seed), device
        self.slot_channel = np.asarray(slot_channel, np.int64)
        self.slots = len(self.slot_channel)
        self.classes = []
        keys = sorted({(self.specs[c].n, self.specs[c].points) for c in self.slot_channel})
        for key in keys:
            idx = np.nonzero([(self.specs[c].n, self.specs[c].points) == key for c in self.slot_channel])[0]
            chans = sorted(set(int(self.slot_channel[i]) for i in idx))
            local = {c: j for j, c in enumerate(chans)}
            tables = DeviceTables([self.specs[c] for c in chans], device)
            B = len(idx)
            ws = Workspace(key[0], key[1], B, mode, device)
            chan = np.array([local[int(self.slot_channel[i])] for i in idx])
            ns = np.array([self.specs[int(self.slot_channel[i])].nscreens for i in idx])
            gb = GraphedBatch(ws, tables, chan, ns, self.seed)
            self.classes.append({"n": key[0], "points": key[1], "slots": idx, "ws": ws, "tables": tables, "gb": gb,
                                 "stream": torch.cuda.Stream(device=device)})   # default priority
        self.out = torch.empty(self.slots, _lib.PA_NUM_MOMENTS, dtype=torch.float64, device=device)
        self._slot_index = [torch.as_tensor(c["slots"], device=device) for c in self.classes]
        self.slot_channel_dev = torch.as_tensor(self.slot_channel.astype(np.int32), device=device)
        self.graph = None

    def ensemble(self, beam_samples: int, pdt_samples: int) -> Ensemble:
        """Device accumulators for this batch's channels (pupils: the largest
        pupil count among them)."""
        npup = {len(s.pupils) for s in self.specs}
        if len(npup) != 1:
            raise ValueError("an ensemble needs the same number of pupils in every channel")
        return Ensemble(len(self.specs), npup.pop(), beam_samples, pdt_samples, self.device)

    def run(self, rid: np.ndarray, ensemble: Optional[Ensemble] = None) -> torch.Tensor:
        rid = np.asarray(rid, np.int64)
        if len(rid) != self.slots:
            raise ValueError("rid length must equal the number of slots")
        if self.graph is None:
            # eager warm-up per class (uploads channel indices and screen counts)
            for c in self.classes:
                gb = c["gb"]
                run_batch(c["ws"], c["tables"], gb.chan, gb.nscreens, seed=self.seed, rid=rid[c["slots"]])
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                main = torch.cuda.current_stream()
                for c in self.classes:
                    c["stream"].wait_stream(main)
                    with torch.cuda.stream(c["stream"]):
                        c["gb"]._capture(overlap=False)
                for c in self.classes:
                    main.wait_stream(c["stream"])
            self.graph = g
            torch.cuda.synchronize()
        for c in self.classes:
            upload(c["ws"].rid[:len(c["slots"])], rid[c["slots"]], np.int64)
        self.graph.replay()
        for c, ix in zip(self.classes, self._slot_index):
            self.out.index_copy_(0, ix, c["ws"].moments[:len(c["slots"])])
        if ensemble is not None:
            ensemble.add(self.out, self.slot_channel_dev)
        return self.out


def default_batch_for(n: int, points: Optional[int] = None, screens: Optional[int] = None,
                      mode: str = "f16x3", device: Optional[torch.device] = None) -> int:
    """Realizations per launch: the autotuner's cached choice for this shape on
    this device when there is one, else about 1 GiB of fields (2^27 complex64
    points)."""
    if points is not None and screens is not None and device is not None:
        from . import autotune
        e = autotune.lookup_batch(n, points, screens, mode, device)
        if e is not None:
            return e["batch"]
    return int(max(1, min(256, (1 << 27) // (n * n))))


_WORKSPACES: Dict[tuple, Workspace] = {}


def get_workspace(n: int, points: int, batch: int, mode: str, device: torch.device) -> Workspace:
    """Cached workspace per (n, points, batch, mode, device)."""
    key = (n, points, batch, mode, device.index or 0)
    ws = _WORKSPACES.get(key)
    if ws is None:
        ws = Workspace(n, points, batch, mode, device)
        _WORKSPACES[key] = ws
    return ws


def release_workspaces() -> None:
    _WORKSPACES.clear()


MOMENT_NAMES = ("eta", "mean_x", "mean_y", "mean_x2", "mean_xy", "mean_y2", "mean_x2_r")


def structure_function(phi: torch.Tensor) -> np.ndarray:
    """Row structure function of float32 device screens phi (n, n) or (B, n, n),
    row-major: D[r - 1] = mean over rows of sum_c (phi[c] - phi[c + r])^2 / (n - r)
    (theory/phase_screens/sf.py:4-8; simulations/structure_function.py:11-12).
    Returns (n - 1,) or (B, n - 1) float64 on the host."""
    if phi.dim() not in (2, 3) or phi.shape[-2] != phi.shape[-1]:
        raise ValueError(f"structure_function takes square screens (n, n) or (B, n, n), got {tuple(phi.shape)}")
    single = phi.dim() == 2
    p = (phi[None] if single else phi).contiguous()
    if p.dtype != torch.float32 or not p.is_cuda:
        raise TypeError("structure_function takes float32 device screens")
    B, n = p.shape[0], p.shape[-1]
    scratch = torch.empty(call("pa_structure_function_scratch", n, B), dtype=torch.uint8, device=p.device)
    out = torch.empty(B, n - 1, dtype=torch.float64, device=p.device)
    call("pa_structure_function", n, ptr(p), ptr(scratch), ptr(out), B, _lib.stream_handle())
    res = out.cpu().numpy()
    return res[0] if single else res


def field_moments(u: torch.Tensor, delta: float, pupils: Sequence[Tuple[float, float, float]] = (),
                  tracked: bool = False) -> np.ndarray:
    """Moments of row-major device fields u (B, N, N) or (N, N) (measures.py:1-38):
    returns (B, PA_NUM_MOMENTS) float64 on the host.  With `tracked` each pupil is
    centred on the field's own centroid (simulations/pdt.py:63-66)."""
    if u.dim() == 2:
        u = u.unsqueeze(0)
    u = u.contiguous()
    if u.dtype != torch.complex64:
        u = u.to(torch.complex64)
    B, n = u.shape[0], u.shape[-1]
    T = call("pa_tile_rows", n)
    dev = u.device
    stream = _lib.stream_handle()
    chp = torch.tensor([[delta, 0.0, 0.0, 0.0]], dtype=torch.float64, device=dev)
    pup_np = np.zeros((1, _lib.PA_MAX_PUPILS, 3))
    for j, p in enumerate(pupils):
        pup_np[0, j] = p
    pup = torch.from_numpy(pup_np).to(dev)
    part = torch.empty(B, n // T, _lib.PA_PART_SUMS, dtype=torch.float64, device=dev)
    out = torch.empty(B, _lib.PA_NUM_MOMENTS, dtype=torch.float64, device=dev)
    chan = torch.zeros(B, dtype=torch.int32, device=dev)
    if tracked:
        call("pa_field_partials", n, ptr(u), ptr(part), ptr(chp), ptr(pup), 0, None, ptr(chan), B, stream)
        call("pa_moments_finalize", n, ptr(part), ptr(chp), 0, ptr(chan), ptr(out), B, stream)
        first = out.clone()
        call("pa_field_partials", n, ptr(u), ptr(part), ptr(chp), ptr(pup), len(pupils), ptr(first),
             ptr(chan), B, stream)
    else:
        call("pa_field_partials", n, ptr(u), ptr(part), ptr(chp), ptr(pup), len(pupils), None, ptr(chan), B,
             stream)
    call("pa_moments_finalize", n, ptr(part), ptr(chp), len(pupils), ptr(chan), ptr(out), B, stream)
    return out.cpu().numpy()
