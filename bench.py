"""Throughput of the split-step path on one MI355X.

Workload: the 2048^2 / 50 km / 5-screen channel (lambda 808 nm, w0 0.12 m,
delta 1.5 mm, MVK Cn2 5e-16, l0 6 mm, L0 1 km, 1024 log-polar spectral points
over [1/15000, 2/6e-3] 1/m, circular pupil r = 0.2 m).  A step is one batch of
realizations through the whole hot path: device spectral draw, screen
synthesis (FP64-reduced plane-wave tables + split-fp16 MFMA GEMM), 6 fused
Fresnel steps (the first, from the separable Gaussian source, as two 1-D
propagations per realization; --no-separable runs it as a 2-D step), 5 screen
multiplies, and the beam-moment / pupil reductions.
`value` is realizations per second of the batched form in the timed region,
where every kernel is launched on the launch stream between HIP events (the
per-kernel averages feed the rooflines); the same batches replayed as a
captured HIP graph, the plain form (batch 1, graph), and the drop-in API used
as pyAtmosphere is used (one Channel.run() and measures.eta per realization,
host draws) are reported beside it.

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
    from pyatmosphere_amd import workloads
    return workloads.channel_2048_50km(CFG["radius"])


def alg_bytes_per_realization(n, S):
    """Bytes the 2-D split-step schedule moves per realization: the first row
    pass writes the field (8 B/element), S+1 column passes and S row passes
    read and write it (16 B), the last row pass reads it (8 B); each screen's
    float32 screen angle is written once by the GEMM and read once by its row
    pass (8 B), which forms the screen factor from it."""
    return n * n * (8 + 16 * (S + 1) + 16 * S + 8 + 8 * S)


def executed_bytes_per_realization(n, S, separable):
    """Field and angle bytes of the schedule actually run: with the separable
    first step there is no first row pass or first column pass, and the first
    screen pass writes the field without reading it (its 1-D vectors are
    O(N))."""
    if not separable:
        return alg_bytes_per_realization(n, S)
    return n * n * (16 * S + 16 * (S - 1) + 8 + 8 + 8 * S)


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


def timed_region(engine, torch, ws, tables, B, steps, warmup, seed):
    """The timed region: `steps` batches of B realizations (fresh realization
    ids each) after `warmup` batches, every kernel launched eagerly on the
    launch stream (torch's current stream) between a pair of HIP events.
    Returns (wall seconds, moments of the last batch, average device ms per
    launch of each kernel family)."""
    from pyatmosphere_amd import _lib
    n, S = ws.n, CFG["S"]
    acc = {}

    def timed(name, fn, record):
        if not record:
            fn()
            return
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        acc.setdefault(name, []).append((s, e))

    def batch(i, record):
        run = engine.StepRunner(ws, tables, np.zeros(B), np.full(B, S), rid=np.arange(i * B, (i + 1) * B))
        st = run.stream
        sep = engine.SEPARABLE_SOURCE
        angs = None
        if engine.SYNTH_WHOLE_BATCH and run.whole_batch_synthesis():
            # all screens' synthesis first: the draw per screen, the tables and
            # the GEMM once over the S x B (screen, realization) pairs
            angs = run.synthesize_all(seed, timed=lambda name, fn: timed(name, fn, record))
        timed("source_fresnel" if sep else "row_first", lambda: run.begin(None, sep), record)
        for k in range(S + 1):
            if not (run.sep0 and k == 0):
                timed("col_step", lambda k=k: run.fresnel(k), record)
            if k < S:
                if angs is None:
                    timed("screen_draw_prep", lambda k=k: engine.draw_prep_device(ws, tables, seed, k, B, None, st),
                          record)
                    timed("screen_tables", lambda k=k: engine.plane_wave_tables(ws, tables, k, B, None, st), record)
                    timed("screen_gemm", lambda k=k: engine.screen_gemm(ws, k, B, None, st, ang=ws.ang), record)
                a = ws.ang if angs is None else angs[k]
                timed("row_mid_source" if (run.sep0 and k == 0) else "row_mid", lambda k=k, a=a: run.apply(k, ang=a),
                      record)
        timed("row_last", lambda: _lib.call(
            "pa_row_last", n, engine.ptr(ws.work), None, engine.ptr(ws.part), engine.ptr(tables.chparam),
            engine.ptr(tables.pup), tables.npup, None, engine.ptr(ws.chan), engine.ptr(run.tw), B, st), record)
        timed("moments_finalize", lambda: _lib.call(
            "pa_moments_finalize", n, engine.ptr(ws.part), engine.ptr(tables.chparam), tables.npup,
            engine.ptr(ws.chan), engine.ptr(ws.moments), B, st), record)
        return ws.moments[:B]

    for i in range(warmup):
        batch(i, False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        m = batch(warmup + i, True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kms = {k: float(np.mean([s.elapsed_time(e) for s, e in v])) for k, v in acc.items()}
    return dt, m, kms


def dropin_rate(torch, spec, realizations=20):
    """Realizations/s of the drop-in API used as pyAtmosphere is used: a
    Channel built from the reference's constructors (the same 2048^2 channel),
    one `channel.run()` per realization with the reference's host draws
    (np.random, reference order) and the pupil transmittance read back as a
    Python float (measures.eta), as the generic Simulation loop does."""
    import pyatmosphere_amd as pa
    ch = pa.Channel(
        grid=pa.RectGrid(resolution=spec.n, delta=spec.delta),
        source=pa.GaussianSource(wvl=spec.wvl, w0=spec.w0, F0=np.inf),
        path=pa.IdenticalPhaseScreensPath(
            phase_screen=pa.SSPhaseScreen(model=pa.MVKModel(Cn2=5e-16, l0=6e-3, L0=1e3),
                                          f_grid=pa.RandLogPolarGrid(points=spec.points, f_min=1 / 15000,
                                                                     f_max=2 / 6e-3)),
            length=spec.length, count=spec.nscreens),
        pupil=pa.CirclePupil(radius=spec.pupils[0][0]))
    np.random.seed(0)
    for _ in range(2):
        pa.measures.eta(ch, output=ch.run())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(realizations):
        pa.measures.eta(ch, output=ch.run())
    torch.cuda.synchronize()
    return realizations / (time.perf_counter() - t0)


def mixed_rate(engine, torch, dev, mode, steps=5, warmup=1):
    """Realizations/s of the mixed-channel workload (256 slots: 512^2, 1024^2 and
    2048^2 grids, 3-10 screens; workloads.mixed_channels): one realization per
    slot per replay of one captured graph, each replay's moments added to the
    per-channel device ensembles (float64 beam sums, 64-bin PDT histograms)."""
    from pyatmosphere_amd import workloads
    specs = workloads.mixed_channels()
    mb = engine.MixedBatch(specs, np.arange(len(specs)), seed=99, mode=mode, device=dev)
    ens = mb.ensemble(beam_samples=2000, pdt_samples=6000)
    S = len(specs)
    for i in range(warmup + 1):
        mb.run(np.arange(S) + i * S)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        m = mb.run(np.arange(S) + (i + warmup + 1) * S, ensemble=ens)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    assert np.all(np.isfinite(m.cpu().numpy()[:, :7])), "non-finite moments"
    assert int(ens.counts[:, 1].min().item()) == steps
    classes = {str(c["n"]): int(len(c["slots"])) for c in mb.classes}
    screens = float(np.mean([s.nscreens for s in specs]))
    del mb, ens
    torch.cuda.empty_cache()
    return {"realizations_per_s": steps * S / dt, "slots": S, "slots_per_grid": classes,
            "mean_screens": screens, "ms_per_replay": dt / steps * 1e3,
            "full_ensembles_s": 6000 * dt / steps}


def calibration(torch, dev):
    """Practical ceilings measured on this device: the library's fp16 GEMM
    (torch.matmul -> hipBLASLt, 8192^3, contiguous operands), and this
    package's own HBM streaming kernels (pa_hbm_stream, 16-byte vectors) over
    1 GiB: copy (read + write bytes), read-only and write-only, each the best
    of 2/4/8 vectors in flight per thread with and without non-temporal hints."""
    from pyatmosphere_amd import _lib

    def ev_ms(fn, reps):
        for _ in range(3):
            fn()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            fn()
        e.record()
        e.synchronize()
        return s.elapsed_time(e) / reps
    n = 8192
    a = torch.randn(n, n, device=dev, dtype=torch.float16)
    b = torch.randn(n, n, device=dev, dtype=torch.float16)
    gemm = 2.0 * n ** 3 / (ev_ms(lambda: torch.matmul(a, b), 10) * 1e-3) / 1e12
    del a, b
    nbytes = 1 << 30
    src = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    sink = torch.zeros(4, dtype=torch.int32, device=dev)
    st = _lib.stream_handle()
    best = {}
    for kind, name, mult in ((0, "copy", 2), (1, "read", 1), (2, "write", 1)):
        for unroll in (2, 4, 8):
            for nt in (0, 1):
                args = {0: (src, dst), 1: (src, sink), 2: (None, dst)}[kind]
                ms = ev_ms(lambda: _lib.call("pa_hbm_stream", kind, _lib.ptr(args[0]), _lib.ptr(args[1]), nbytes,
                                             unroll, nt, st), 10)
                gbs = mult * nbytes / (ms * 1e-3) / 1e9
                if gbs > best.get(name, (0.0,))[0]:
                    best[name] = (gbs, unroll, nt)
    torch_copy = 2.0 * nbytes / (ev_ms(lambda: dst.copy_(src), 10) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return {"fp16_gemm_tflops": gemm, "fp16_gemm": "torch.matmul fp16 8192^3 (hipBLASLt)",
            "copy_gbs": best["copy"][0], "read_gbs": best["read"][0], "write_gbs": best["write"][0],
            "stream_shapes": {k: {"unroll": v[1], "nontemporal": v[2]} for k, v in best.items()},
            "torch_copy_gbs": torch_copy,
            "stream": "pa_hbm_stream over 1 GiB, 16-byte vectors; copy counts read + write bytes"}


def profile_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary of
    this workload (FETCH_SIZE and WRITE_SIZE, calibrated against a plain copy),
    or None when no summary carries it."""
    path = os.path.join(ROOT, "profiles", f"{workload}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic-json:"):
            d = json.loads(line.split(":", 1)[1])
            v = d.get(kernel)
            return None if v is None else float(v)
    return None


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
    (run before the GPU is touched), on bounded samples of two channels: the
    2048^2 / 50 km / 5-screen channel and the 512^2 / 10 km / 5-screen channel
    (seed 0).  Each as one process with BLAS on every core (numpy FFT/exp
    single-threaded, as in the reference) and as one single-threaded process
    per core (at most 16, the box's CPU share); `value` is the better 2048^2
    rate."""
    from oracle import cpu_bench
    from pyatmosphere_amd import workloads
    procs = min(cpu_bench.cores(), 16)
    c2 = cpu_bench.spec_arrays(workloads.channel_2048_50km(), 0.2)
    c1 = cpu_bench.spec_arrays(workloads.channel_512_10km(), 0.12)
    n_s, t_s = cpu_bench.single(c2, seconds_target, 8)
    n_p, t_p = cpu_bench.per_core(c2, procs, 1)
    m_s, u_s = cpu_bench.single(c1, 10.0, 100)
    m_p, u_p = cpu_bench.per_core(c1, procs, 4)
    v_s, v_p = n_s / t_s, n_p / t_p
    best = max(v_s, v_p)
    threads = int(os.environ.get("OPENBLAS_NUM_THREADS", os.environ.get("OMP_NUM_THREADS", procs)))
    return {"value": best, "unit": "realizations/s", "cores": procs if v_p >= v_s else threads, "kind": "port",
            "sample": (f"2048^2 channel: {n_s} realizations in one process ({t_s:.1f} s, BLAS threads {threads}); "
                       f"{n_p} realizations as {procs} single-threaded processes ({t_p:.1f} s wall); "
                       "numpy complex64 op-for-op port"),
            "cpu_model": cpu_bench.cpu_model(), "nproc": os.cpu_count(), "affinity_cores": cpu_bench.cores(),
            "variants": {"single_process": v_s, "process_per_core": v_p},
            "channel_512": {"single_process": m_s / u_s, "process_per_core": m_p / u_p,
                            "sample": f"{m_s} realizations in one process ({u_s:.1f} s); {m_p} as {procs} "
                                      f"processes ({u_p:.1f} s wall)"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--mode", default="f16x3")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--eager", action="store_true", help="skip the graph-replay timing of the batched form")
    ap.add_argument("--no-overlap", action="store_true", help="run screen synthesis on the main stream")
    ap.add_argument("--no-separable", action="store_true",
                    help="take the source's first Fresnel step as a 2-D step (row_first + col_step)")
    ap.add_argument("--per-screen-synthesis", action="store_true",
                    help="synthesize the screens one at a time instead of all screens of a batch at once")
    ap.add_argument("--no-tune", action="store_true", help="library default launch shapes, no autotuning")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-channel (256-slot) workload")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 512 MiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the path shards realizations across GPUs with one final "
              "all-reduce of the moment sums (not part of this build)", file=sys.stderr)
        sys.exit(2)
    # the host baseline runs first, before anything touches the GPU (its worker
    # processes are started from a process with no device context)
    cpu = None if args.no_cpu else cpu_baseline()
    import torch
    from pyatmosphere_amd import engine
    dev = torch.device("cuda", 0)
    spec = channel_spec(engine)
    tables = engine.DeviceTables([spec], dev)
    seed = 20241202

    if args.no_overlap:
        engine.OVERLAP_DEFAULT = False
    if args.no_separable:
        engine.SEPARABLE_SOURCE = False
    if args.per_screen_synthesis:
        engine.SYNTH_WHOLE_BATCH = False
    if args.copy_calibration:
        # three copies of 512 MiB by the package's streaming kernel (kind 0, 4
        # vectors per thread): a dispatch of known bytes for scaling the
        # FETCH_SIZE / WRITE_SIZE counters of a profiled run
        from pyatmosphere_amd import _lib
        src = torch.empty(512 * 2 ** 20 // 4, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            _lib.call("pa_hbm_stream", 0, _lib.ptr(src), _lib.ptr(dst), 512 * 2 ** 20, 4, 0, _lib.stream_handle())
        torch.cuda.synchronize()
        del src, dst

    from pyatmosphere_amd import autotune
    launch = {}

    def tuned(B):
        # screen launch shapes for this batch, measured now (a fraction of a second)
        if args.no_tune:
            return
        e = autotune.tune_screen(spec.n, spec.points, B, args.mode, dev, spec=spec)
        launch[f"batch{B}"] = {k: e[k] for k in ("gemm_tile", "table_rows", "col_group", "gemm_ms", "tables_ms",
                                                 "col_ms")}

    plain = None
    if not args.skip_plain:
        # plain: batch 1
        tuned(1)
        ws1 = engine.Workspace(spec.n, spec.points, 1, args.mode, dev)
        dt1, _ = time_batches(engine, torch, ws1, tables, 1, max(args.steps, 5), args.warmup, seed,
                              graph=not args.eager)
        plain = max(args.steps, 5) / dt1
        del ws1
        torch.cuda.empty_cache()

    dropin = None if args.skip_plain else dropin_rate(torch, spec)

    B = args.batch
    tuned(B)
    ws = engine.Workspace(spec.n, spec.points, B, args.mode, dev)
    graph_rate = None
    if not args.eager:
        dtg, _ = time_batches(engine, torch, ws, tables, B, args.steps, args.warmup, seed, graph=True)
        graph_rate = args.steps * B / dtg
    dt, mom, kt = timed_region(engine, torch, ws, tables, B, args.steps, args.warmup, seed)
    batched = args.steps * B / dt
    m = mom.cpu().numpy()
    assert np.all(np.isfinite(m[:, :7])), "non-finite moments"

    S, n, P = CFG["S"], CFG["n"], CFG["P"]
    workload = f"splitstep_{n}_b{B}"
    # dominant kernel: the screen GEMM (MFMA-bound).  Algorithmic flops are the
    # useful flops of the screen contraction phi = Re (a E_y) E_x, 4 N^2 P per
    # screen and realization (real output, K = 2P real columns): `achieved`.
    # The GEMM executes fewer columns, K = Kp - kbeg (two per direct spectral
    # point plus the 32 of the polynomial block that carries the low-frequency
    # points; the mean over the last batch), reported under `executed`, with
    # the three split-fp16 products counted in `issued_frac`.
    # With whole-batch synthesis one tables launch and one GEMM launch cover
    # all S x B (screen, realization) pairs (G screens per launch); the draw
    # runs per screen (B per launch).
    whole = engine.SYNTH_WHOLE_BATCH and S * B * engine.synth_bytes_per_element(n, P, ws.mode) <= \
        engine.SYNTH_BATCH_MAX_BYTES
    G = S * B if whole else B
    sp = ws.synth_buffers(S).sparam[:G] if whole else ws.sparam[:B]
    Kp = ((2 * P + 31) // 32) * 32
    kbeg = sp[:, 4].cpu().numpy()
    k_eff = float(np.mean(Kp - kbeg))
    t_gemm = kt["screen_gemm"] * 1e-3
    alg_tfs = G * 4.0 * n * n * P / t_gemm / 1e12
    gemm_tfs = G * 2.0 * n * n * k_eff / t_gemm / 1e12
    issued = 1.0 if args.mode == "f32" else 3.0
    peak = 157.3 if args.mode == "f32" else F16_PEAK_TFLOPS
    # the chain's time per B screens
    t_chain = ((kt["screen_gemm"] + kt["screen_tables"]) * B / G + kt["screen_draw_prep"]) * 1e-3
    # HBM-bound FFT passes: algorithmic bytes per element (read + write the
    # field; the row pass also reads the float32 screen angle)
    # (bytes per element, the streaming ceiling of the same direction)
    passes = {"col_step": (16, "copy_gbs"), "row_mid": (20, "copy_gbs"), "row_first": (8, "write_gbs"),
              "row_mid_source": (12, "copy_gbs"), "row_last": (8, "read_gbs")}
    cal = calibration(torch, dev)
    hbm = {}
    for k, (per, ceil) in passes.items():
        if k not in kt:
            continue
        gbs = B * n * n * per / (kt[k] * 1e-3) / 1e9
        hbm[k] = {"bytes_per_element": per, "achieved": gbs, "frac": gbs / HBM_PEAK_GBS,
                  "ceiling": ceil, "frac_of_ceiling": gbs / cal[ceil], "traffic": profile_traffic(k, workload)}
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
        "dtype": "f32 (fields, FFTs), split-f16 MFMA with f32 accumulation (screens), f64 (tables, moments)",
        "data": "synthetic (device Philox spectral draws of the reference's distributions)",
        "config": {"workload": "2048^2 grid, 50 km, 5 SS/MVK screens, P=1024, batched", "batch": B,
                   "gemm_mode": args.mode},
        "gemm_k": {"Kp": Kp, "k_mean": k_eff, "k_low_mean": float(sp[:, 5].mean().item()),
                   "screens_per_launch": G},
        "roofline": {"bound": "mfma", "kernel": "screen_gemm", "achieved": alg_tfs, "peak": peak,
                     "unit": "TFLOP/s", "frac": alg_tfs / peak, "traffic": profile_traffic("screen_gemm", workload),
                     "flops": "4 N^2 P per screen (useful flops of the screen contraction), "
                              "gemm_k.screens_per_launch screens per launch",
                     "executed": {"k": k_eff, "achieved": gemm_tfs, "frac": gemm_tfs / peak,
                                  "issued_frac": issued * gemm_tfs / peak,
                                  "issued_frac_of_library": None if args.mode == "f32"
                                  else issued * gemm_tfs / cal["fp16_gemm_tflops"]},
                     "synthesis_chain": {"kernels": "draw_prep + tables (incl. polynomial block) + gemm",
                                         "achieved": B * 4.0 * n * n * P / t_chain / 1e12,
                                         "frac": B * 4.0 * n * n * P / t_chain / 1e12 / peak}},
        "calibration": cal,
        "roofline_hbm": hbm,
        "plain_realizations_per_s": plain,
        "dropin_realizations_per_s": dropin,
        "graph_realizations_per_s": graph_rate,
        "kernel_ms": kt,
        "launch_shapes": launch,
        "alg_bytes_per_realization": alg_bytes_per_realization(n, S),
        "executed_bytes_per_realization": executed_bytes_per_realization(n, S, engine.SEPARABLE_SOURCE),
        "separable_first_step": engine.SEPARABLE_SOURCE,
    }
    del ws
    torch.cuda.empty_cache()
    out["mixed_channels"] = None if args.no_mixed else mixed_rate(engine, torch, dev, args.mode)
    out["surrogate_pdt"] = None if args.no_surrogate else surrogate_check(torch, dev)
    if cpu is not None:
        out["cpu_baseline"] = cpu
        out["cpu_baseline"]["gpu_over_cpu"] = batched / cpu["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
