"""Launch-shape autotuner for the split-step path.

The results of the kernels do not depend on their launch shapes, only their
speed does, so the shapes are chosen by measurement on the device at hand:

* per (grid n, spectral points P, batch B, GEMM mode): the screen GEMM's
  block tile (256 x 256 with 8 waves, or 128 x 128 with 4 waves), the rows
  per plane-wave-table block (4, 8, 16 or 32) and the columns per block of
  the column FFT pass (pa_col_groups: the layout's default, half or twice
  it), each kernel timed alone with HIP events on the same inputs
  (`tune_screen`);
* per (n, P, screens, mode): the batch size and whether screen synthesis runs
  on a second stream, by timing whole graph-replayed batches (`tune_batch`).

Choices are kept in a JSON cache ($PA_TUNE_CACHE, else
~/.cache/pyatmosphere_amd/tune.json) keyed by the device's name and
architecture, the library version and the problem shape.  Workspaces pick up
a cached screen shape when they are created; without an entry the library's
own defaults apply (pa_screen_gemm / pa_screen_tables with shape 0).
"""
import json
import os
import tempfile
import time
from typing import Any, Callable, Dict, Iterable, Optional

import numpy as np
import torch

GEMM_TILES = (1, 2)
TABLE_ROWS = (4, 8, 16, 32)
_DEVICE_NAMES: Dict[int, str] = {}


def cache_path() -> str:
    p = os.environ.get("PA_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "pyatmosphere_amd", "tune.json")


def _load() -> Dict[str, Any]:
    try:
        with open(cache_path()) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(key: str, entry: Dict[str, Any]) -> None:
    """Merge one entry into the cache file (atomic replace)."""
    d = _load()
    d[key] = entry
    path = cache_path()
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=os.path.dirname(path) or ".", suffix=".tmp")
    with os.fdopen(fd, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def device_name(device: torch.device) -> str:
    idx = device.index or 0
    if idx not in _DEVICE_NAMES:
        p = torch.cuda.get_device_properties(idx)
        _DEVICE_NAMES[idx] = f"{p.name}|{getattr(p, 'gcnArchName', '')}|{p.multi_processor_count}cu"
    return _DEVICE_NAMES[idx]


def _prefix(device: torch.device) -> str:
    from . import _lib
    return f"{device_name(device)}|lib{_lib.call('pa_version')}"


def screen_key(n: int, points: int, batch: int, mode: str, device: torch.device) -> str:
    return f"{_prefix(device)}|screen|n{n}|P{points}|B{batch}|{mode}"


def batch_key(n: int, points: int, screens: int, mode: str, device: torch.device) -> str:
    return f"{_prefix(device)}|batch|n{n}|P{points}|S{screens}|{mode}"


def col_groups(n: int):
    """Columns per column-pass block that pa_col_step takes for grid size n."""
    import ctypes
    from . import _lib
    out = (ctypes.c_int * 3)()
    k = _lib.call("pa_col_groups", n, ctypes.cast(out, ctypes.c_void_p))
    return tuple(out[:k])


def lookup(n: int, points: int, batch: int, mode: str, device: torch.device) -> Optional[Dict[str, Any]]:
    """Cached screen launch shape {"gemm_tile", "table_rows", ...} or None."""
    e = _load().get(screen_key(n, points, batch, mode, device))
    if not e or e.get("gemm_tile") not in GEMM_TILES or e.get("table_rows") not in TABLE_ROWS:
        return None
    return e


def lookup_batch(n: int, points: int, screens: int, mode: str, device: torch.device) -> Optional[Dict[str, Any]]:
    """Cached batch choice {"batch", "overlap", ...} or None."""
    e = _load().get(batch_key(n, points, screens, mode, device))
    if not e or not isinstance(e.get("batch"), int) or e["batch"] < 1:
        return None
    return e


def pick(timings: Dict[Any, float]) -> Any:
    """The candidate with the smallest time (the first one listed on ties)."""
    if not timings:
        raise ValueError("no timings")
    best = None
    for k, t in timings.items():
        if best is None or t < timings[best]:
            best = k
    return best


def time_ms(fn: Callable[[], None], reps: int = 5, warmup: int = 1) -> float:
    """Median device time of fn() in ms, HIP events on the current stream."""
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        e.synchronize()
        ts.append(s.elapsed_time(e))
    return float(np.median(ts))


def synthetic_spec(n: int, points: int, screens: int = 5):
    """A channel of the given shape (2048^2-like physics scaled to n) for timing."""
    from . import engine
    delta, wvl, L = 1.5e-3, 808e-9, 50e3
    base = engine.log_polar_base(points, 1 / 15000, 2 / 6e-3)
    psd = engine.annulus_psd(base, 2 * np.pi / wvl, L / screens,
                             lambda k: engine.mvk_psd_n(k, 5e-16, 6e-3, 1e3)).astype(np.float32)
    pos = (np.arange(screens) + 0.5) * L / screens
    return engine.ChannelSpec(n=n, delta=delta, wvl=wvl, w0=0.12, F0=np.inf, length=L, positions=pos,
                              points=points, base=base, psd=psd, pupils=[(0.2, 0.0, 0.0)])


def tune_screen(n: int, points: int, batch: int, mode: str, device: torch.device, reps: int = 5,
                store: bool = True, spec=None) -> Dict[str, Any]:
    """Time every GEMM tile and table-rows candidate on one screen of `batch`
    realizations and cache the fastest of each."""
    from . import engine
    spec = spec or synthetic_spec(n, points)
    tables = engine.DeviceTables([spec], device)
    ws = engine.Workspace(n, points, batch, mode, device)
    ws.chan[:batch].zero_()
    ws.rid[:batch].copy_(torch.arange(batch, dtype=torch.int64))
    st = torch.cuda.current_stream(device).cuda_stream
    engine.draw_device(ws, tables, 12345, 0, batch, st)
    engine.prep_screen(ws, tables, 0, batch, None, st)
    rows_ms = {r: time_ms(lambda r=r: engine.plane_wave_tables(ws, tables, 0, batch, None, st, rows=r), reps)
               for r in TABLE_ROWS}
    tiles = GEMM_TILES if n >= 512 else (2,)
    tile_ms = {t: time_ms(lambda t=t: engine.screen_gemm(ws, 0, batch, None, st, ang=ws.ang, tile=t), reps)
               for t in tiles}
    ws.work[:batch].zero_()
    tw = engine.twiddles(n, device)
    col_ms = {c: time_ms(lambda c=c: engine.call("pa_col_step", n, engine.ptr(ws.work), engine.ptr(tables.htab),
                                                  engine.ptr(ws.chan), None, 0, tables.max_steps, engine.ptr(tw), c,
                                                  batch, st), reps)
              for c in col_groups(n)}
    entry = {"gemm_tile": pick(tile_ms), "table_rows": pick(rows_ms), "col_group": pick(col_ms),
             "gemm_ms": {str(k): v for k, v in tile_ms.items()},
             "tables_ms": {str(k): v for k, v in rows_ms.items()},
             "col_ms": {str(k): v for k, v in col_ms.items()}, "time": time.time()}
    del ws
    if store:
        _store(screen_key(n, points, batch, mode, device), entry)
    return entry


def tune_batch(n: int, points: int, screens: int, mode: str, device: torch.device,
               candidates: Iterable[int] = (1, 2, 4, 8, 16, 32, 64), steps: int = 3,
               max_bytes: Optional[int] = None, store: bool = True, spec=None) -> Dict[str, Any]:
    """Realizations per second of graph-replayed batches for each batch size
    (screen shapes tuned first) with and without the second stream; caches the
    fastest (batch, overlap)."""
    from . import engine
    spec = spec or synthetic_spec(n, points, screens)
    tables = engine.DeviceTables([spec], device)
    if max_bytes is None:
        free, _ = torch.cuda.mem_get_info(device)
        max_bytes = free // 2
    rates: Dict[tuple, float] = {}
    for B in candidates:
        # field (8 B), phi2 per screen (4 B) and the operands (~8 B at K = 2P) per element
        if B * n * n * (8 + 4 * (1 + screens) + 8) > max_bytes:
            continue
        tune_screen(n, points, B, mode, device, store=store, spec=spec)
        ws = engine.Workspace(n, points, B, mode, device)
        for ov in (False, True):
            saved = engine.OVERLAP_RULE
            engine.OVERLAP_RULE = lambda b, nn, ov=ov: ov
            try:
                gb = engine.GraphedBatch(ws, tables, np.zeros(B), np.full(B, screens), 777)
                gb.run(np.arange(B))
                gb.run(np.arange(B, 2 * B))
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for i in range(steps):
                    gb.run(np.arange((i + 2) * B, (i + 3) * B))
                torch.cuda.synchronize()
                rates[(B, ov)] = steps * B / (time.perf_counter() - t0)
                del gb
            finally:
                engine.OVERLAP_RULE = saved
        del ws
        engine.release_workspaces()
        torch.cuda.empty_cache()
    if not rates:
        raise RuntimeError("no batch size fits the memory limit")
    best = pick({k: -v for k, v in rates.items()})
    entry = {"batch": int(best[0]), "overlap": bool(best[1]),
             "rates": {f"{b}{'+overlap' if o else ''}": r for (b, o), r in rates.items()}, "time": time.time()}
    if store:
        _store(batch_key(n, points, screens, mode, device), entry)
    return entry
