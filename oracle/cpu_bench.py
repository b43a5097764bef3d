"""Host timing of the CPU baseline: the op-for-op complex64 numpy port of the
reference's path (refport_c64) on a bounded sample of realizations, as one
process with BLAS on every core (numpy's FFT and exp single-threaded, as in
the reference) and as one single-threaded process per core.

TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg runs it, before the
GPU is touched; worker processes import numpy only.

    python -m oracle.cpu_bench --worker SPEC.npz COUNT SEED   (one worker)
"""
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

from oracle import refport_c64

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _realize(rs, p):
    return refport_c64.realization(rs, int(p["n"]), float(p["delta"]), float(p["wvl"]), float(p["w0"]),
                                   float(p["length"]), p["positions"], p["base"], p["psd"], float(p["radius"]))


def single(p, seconds: float, max_count: int, seed: int = 0):
    """Sequential realizations in this process until `seconds` or
    `max_count`; returns (count, elapsed)."""
    rs = np.random.RandomState(seed)
    t0 = time.perf_counter()
    count = 0
    while True:
        _realize(rs, p)
        count += 1
        el = time.perf_counter() - t0
        if el >= seconds or count >= max_count:
            return count, el


def per_core(p, procs: int, per_proc: int, seed: int = 0):
    """`procs` single-threaded worker processes, `per_proc` realizations each,
    started together; returns (total realizations, wall seconds)."""
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "spec.npz")
        np.savez(path, **p)
        env = dict(os.environ, OPENBLAS_NUM_THREADS="1", OMP_NUM_THREADS="1", MKL_NUM_THREADS="1",
                   PYTHONPATH=ROOT + os.pathsep + os.environ.get("PYTHONPATH", ""))
        t0 = time.perf_counter()
        workers = [subprocess.Popen([sys.executable, "-m", "oracle.cpu_bench", "--worker", path, str(per_proc),
                                     str(seed + 1 + i)], env=env, cwd=ROOT) for i in range(procs)]
        codes = [w.wait() for w in workers]
        wall = time.perf_counter() - t0
    if any(codes):
        raise RuntimeError(f"cpu baseline worker failed: {codes}")
    return procs * per_proc, wall


def spec_arrays(spec, radius):
    return {"n": spec.n, "delta": spec.delta, "wvl": spec.wvl, "w0": spec.w0, "length": spec.length,
            "positions": np.asarray(spec.positions, np.float64), "base": spec.base, "psd": spec.psd,
            "radius": radius}


def _worker(path, count, seed):
    with np.load(path) as z:
        p = {k: z[k] for k in z.files}
    rs = np.random.RandomState(seed)
    for _ in range(count):
        _realize(rs, p)


if __name__ == "__main__":
    if len(sys.argv) == 5 and sys.argv[1] == "--worker":
        _worker(sys.argv[2], int(sys.argv[3]), int(sys.argv[4]))
    else:
        print(__doc__)
