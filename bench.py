"""Throughput of the split-step path on one MI355X.

Workload: the 2048^2 / 50 km / 5-screen channel (lambda 808 nm, w0 0.12 m,
delta 1.5 mm, MVK Cn2 5e-16, l0 6 mm, L0 1 km, 1024 log-polar spectral points
over [1/15000, 2/6e-3] 1/m, circular pupil r = 0.2 m).  A step is one batch of
realizations through the whole hot path: device spectral draw, screen
synthesis (FP64-reduced plane-wave tables + split-fp16 MFMA GEMM), 6 fused
Fresnel steps, 5 screen multiplies, and the beam-moment / pupil reductions.
`value` is realizations per second of the batched form; the plain form
(batch 1) is reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CFG = dict(n=2048, delta=1.5e-3, wvl=808e-9, w0=0.12, L=50e3, S=5, P=1024, Cn2=5e-16, l0=6e-3, L0=1e3,
           f_min=1 / 15000, f_max=2 / 6e-3, radius=0.2)
HBM_PEAK_GBS = 8000.0
F16_PEAK_TFLOPS = 2500.0


def channel_spec(engine):
    c = CFG
    base = engine.log_polar_base(c["P"], c["f_min"], c["f_max"])
    psd = engine.annulus_psd(base, 2 * np.pi / c["wvl"], c["L"] / c["S"],
                             lambda k: engine.mvk_psd_n(k, c["Cn2"], c["l0"], c["L0"])).astype(np.float32)
    pos = (np.arange(c["S"]) + 0.5) * c["L"] / c["S"]
    return engine.ChannelSpec(n=c["n"], delta=c["delta"], wvl=c["wvl"], w0=c["w0"], F0=np.inf, length=c["L"],
                              positions=pos, points=c["P"], base=base, psd=psd,
                              pupils=[(c["radius"], 0.0, 0.0)])


def alg_bytes_per_realization(n, S):
    """Minimal fused schedule: 1 forward-row pass (write), S+1 column passes,
    S row passes with the screen read, 1 final row pass (read), screen write+read."""
    return n * n * (16 * (2 * S + 3) + 8 * S)


def time_batches(engine, torch, ws, tables, B, steps, warmup, seed, graph=True):
    """Time `steps` batches of B realizations (fresh realization ids each) after
    `warmup` batches.  With `graph` the batch's launch sequence is a captured
    HIP graph replayed per step."""
    chan = np.zeros(B)
    ns = np.full(B, CFG["S"])
    gb = engine.GraphedBatch(ws, tables, chan, ns, seed) if graph else None

    def one(i):
        rid = np.arange(i * B, (i + 1) * B)
        if gb is not None:
            return gb.run(rid)
        return engine.run_batch(ws, tables, chan, ns, seed=seed, rid=rid)

    for i in range(max(warmup, 1 if graph else 0)):
        one(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        m = one(warmup + 1 + i)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return dt, m


def kernel_times(engine, torch, ws, tables, B, seed, reps=3):
    """Average device time of each kernel family of one batch, with HIP events
    on the launching stream (torch's current stream)."""
    from pyatmosphere_amd import _lib
    n = ws.n
    run = engine.StepRunner(ws, tables, np.zeros(B), np.full(B, CFG["S"]), rid=np.arange(B))
    acc = {}

    def timed(name, fn):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        acc.setdefault(name, []).append((s, e))

    for _ in range(reps):
        timed("row_first", lambda: run.begin(None))
        for i in range(CFG["S"] + 1):
            timed("col_step", lambda i=i: run.fresnel(i))
            if i < CFG["S"]:
                timed("spectral_draw", lambda i=i: engine.draw_device(ws, tables, seed, i, B, run.stream))
                timed("screen_prep", lambda i=i: _lib.call("pa_screen_prep", ws.points, engine.ptr(ws.rho),
                                                           engine.ptr(ws.theta), engine.ptr(ws.value),
                                                           engine.ptr(ws.fxy), engine.ptr(ws.sparam), ws.mode,
                                                           None, i, B, run.stream))
                timed("screen_tables", lambda i=i: _lib.call(
                    "pa_screen_tables", n, ws.points, engine.ptr(ws.fxy), engine.ptr(ws.value),
                    engine.ptr(ws.sparam), engine.ptr(tables.chparam), engine.ptr(ws.chan), engine.ptr(ws.opa),
                    engine.ptr(ws.opb), engine.ptr(ws.hg), engine.ptr(ws.esc), ws.mode, None, i, B, run.stream))
                timed("screen_gemm", lambda i=i: _lib.call(
                    "pa_screen_gemm", n, ws.points, engine.ptr(ws.opa), engine.ptr(ws.opb), engine.ptr(ws.sparam),
                    engine.ptr(ws.esc), None, engine.ptr(ws.fac), ws.mode, None, i, B, run.stream))
                timed("row_mid", lambda i=i: _lib.call(
                    "pa_row_mid", n, engine.ptr(ws.work), engine.ptr(ws.fac), None, i, engine.ptr(run.tw), B,
                    run.stream))
        timed("row_last", lambda: run.finish(None, True))
    torch.cuda.synchronize()
    return {k: float(np.mean([s.elapsed_time(e) for s, e in v])) for k, v in acc.items()}


def surrogate_check(torch, dev, channels=4, realizations=2000, seed=777):
    """KS distance between the committed surrogate ensemble's PDTs and PDTs
    simulated now for a few fresh parameter draws (8 apertures each)."""
    import glob
    from pyatmosphere_amd.surrogate import PDTSurrogate, SurrogateEnsemble, ks_binned
    from pyatmosphere_amd.surrogate import data as sdata
    files = sorted(glob.glob(os.path.join(ROOT, "pyatmosphere_amd", "surrogate", "weights", "pdt_member*.npz")))
    if not files:
        return None
    ens = SurrogateEnsemble([PDTSurrogate.load(f, device=dev) for f in files])
    ch = sdata.sample_channels(channels, seed=seed)
    t0 = time.perf_counter()
    sim = sdata.simulate(ch, realizations, seed=seed, device=dev, channel_ids=range(5 * 10 ** 6, 5 * 10 ** 6 + channels))
    torch.cuda.synchronize()
    t_sim = time.perf_counter() - t0
    x, te, _ = sdata.dataset(ch, sim)
    t0 = time.perf_counter()
    p, _, spread, fallback = ens.pdt(x)
    t_inf = time.perf_counter() - t0
    ks = ks_binned(p, te)
    return {"ks_mean": float(ks.mean()), "ks_median": float(np.median(ks)),
            "ks_accepted_mean": float(ks[~fallback].mean()) if np.any(~fallback) else None,
            "fallback_fraction": float(fallback.mean()), "pairs": int(len(ks)),
            "realizations_per_pdt": realizations, "surrogate_s": t_inf, "simulation_s": t_sim}


def cpu_baseline(seconds_target=20.0):
    """The op-for-op complex64 numpy port of the reference timed on the host
    on a bounded sample of realizations of the same channel (single process,
    numpy FFT/exp single-threaded, BLAS threads = all cores)."""
    from oracle import refport_c64
    from pyatmosphere_amd import engine
    spec = channel_spec(engine)
    rs = np.random.RandomState(0)
    t0 = time.perf_counter()
    count = 0
    while True:
        refport_c64.realization(rs, spec.n, spec.delta, spec.wvl, spec.w0, spec.length, spec.positions,
                                spec.base, spec.psd, CFG["radius"])
        count += 1
        el = time.perf_counter() - t0
        if el > seconds_target or count >= 8:
            break
    threads = int(os.environ.get("OPENBLAS_NUM_THREADS", os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    return {"value": count / el, "unit": "realizations/s", "cores": threads, "kind": "port",
            "sample": f"{count} realizations of the 2048^2 5-screen channel, single process, "
                      f"numpy complex64 op-for-op port, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--mode", default="f16x3")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--eager", action="store_true", help="launch kernels one by one instead of graph replay")
    ap.add_argument("--no-overlap", action="store_true", help="run screen synthesis on the main stream")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 512 MiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the path shards realizations across GPUs with one final "
              "all-reduce of the moment sums (not part of this build)", file=sys.stderr)
        sys.exit(2)
    import torch
    from pyatmosphere_amd import engine
    dev = torch.device("cuda", 0)
    spec = channel_spec(engine)
    tables = engine.DeviceTables([spec], dev)
    seed = 20241202

    if args.no_overlap:
        engine.OVERLAP_DEFAULT = False
    if args.copy_calibration:
        src = torch.empty(512 * 2 ** 20 // 4, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    plain = None
    if not args.skip_plain:
        # plain: batch 1
        ws1 = engine.Workspace(spec.n, spec.points, 1, args.mode, dev)
        dt1, _ = time_batches(engine, torch, ws1, tables, 1, max(args.steps, 5), args.warmup, seed,
                              graph=not args.eager)
        plain = max(args.steps, 5) / dt1
        del ws1
        torch.cuda.empty_cache()

    B = args.batch
    ws = engine.Workspace(spec.n, spec.points, B, args.mode, dev)
    dt, mom = time_batches(engine, torch, ws, tables, B, args.steps, args.warmup, seed, graph=not args.eager)
    batched = args.steps * B / dt
    kt = kernel_times(engine, torch, ws, tables, B, seed)
    m = mom.cpu().numpy()
    assert np.all(np.isfinite(m[:, :7])), "non-finite moments"

    S, n = CFG["S"], CFG["n"]
    # dominant memory-bound kernel: the fused row pass (inverse FFT, screen, forward FFT)
    row_mid_bytes = B * n * n * (8 + 8 + 4)          # read field, write field, read phi2
    row_mid_gbs = row_mid_bytes / (kt["row_mid"] * 1e-3) / 1e9
    gemm_flops = B * 4.0 * n * n * CFG["P"]           # useful: real output, K = 2P
    gemm_tfs = gemm_flops / (kt["screen_gemm"] * 1e-3) / 1e12
    out = {
        "metric": "realizations/s, 2048^2 5-screen channel (batched)",
        "value": batched,
        "unit": "realizations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 fields, fp64 reductions, split-fp16 MFMA screens",
        "data": "synthetic (device Philox spectral draws)",
        "config": {"workload": "2048^2 grid, 50 km, 5 SS/MVK screens, P=1024, batched", "batch": B,
                   "gemm_mode": args.mode},
        "plain_realizations_per_s": plain,
        "kernel_ms": kt,
        "roofline": {"bound": "hbm", "kernel": "row_mid", "achieved": row_mid_gbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": row_mid_gbs / HBM_PEAK_GBS, "traffic": None},
        "mfma": {"kernel": "screen_gemm", "achieved_useful_tflops": gemm_tfs,
                 "issued_tflops": 3 * gemm_tfs if args.mode != "f32" else gemm_tfs,
                 "peak": F16_PEAK_TFLOPS if args.mode != "f32" else 157.3},
        "alg_bytes_per_realization": alg_bytes_per_realization(n, S),
    }
    out["surrogate_pdt"] = None if args.no_surrogate else surrogate_check(torch, dev)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
