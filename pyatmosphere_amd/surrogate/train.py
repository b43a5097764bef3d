"""Train and evaluate the PDT surrogate on the device.

    python -m pyatmosphere_amd.surrogate.train --grid 1024 --train-channels 3584 --train-real 2000 \
        --test-channels 512 --test-real 2000 --members 0 --chunk 512 --save-data sur_a.npz
    python -m pyatmosphere_amd.surrogate.train --grid 1024 --train-channels 4096 --train-real 2000 \
        --test-channels 0 --members 0 --chunk 512 --first-id 3584 --seed 3024 --save-data sur_b.npz
    (and likewise --first-id 7680 --seed 4024 --save-data sur_c.npz), then
    python -m pyatmosphere_amd.surrogate.train --load-data sur_a.npz,sur_b.npz,sur_c.npz --members 3 \
        --steps 40000 --batch 1024 --width 384 --out pyatmosphere_amd/surrogate/weights

Steps: sample channel parameters over the training domain, simulate them with
the batched engine (8 apertures per channel), train `members` networks with
different seeds, then evaluate on held-out channels simulated with more
realizations: KS distance between the surrogate PDT and the simulated one,
against two references that need no network (the global-mean PDT of the
training set and the PDT of the nearest training channel in the encoded
parameter space), the sampling noise floor of the simulated PDT, and the
time per channel of the surrogate versus the simulation.  Writes the member
weights (.npz) and report.json.
"""
import argparse
import json
import os
import time

import numpy as np
import torch

from . import data as sdata
from .model import NB, PDTSurrogate, SurrogateEnsemble, ks_binned


def noise_floor(te: np.ndarray, realizations: int, rng, reps: int = 4) -> float:
    """Mean KS between a binned PDT and multinomial resamples of it with the
    same number of realizations (what a perfect model would score)."""
    ks = []
    for _ in range(reps):
        s = np.stack([rng.multinomial(realizations, p / p.sum()) / realizations for p in te])
        ks.append(ks_binned(s, te))
    return float(np.mean(ks))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-channels", type=int, default=1500)
    ap.add_argument("--train-real", type=int, default=256)
    ap.add_argument("--test-channels", type=int, default=64)
    ap.add_argument("--test-real", type=int, default=2000)
    ap.add_argument("--members", type=int, default=3)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--width", type=int, default=256)
    ap.add_argument("--out", default="")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--chunk", type=int, default=2000, help="channels simulated per progress line")
    ap.add_argument("--first-id", type=int, default=0, help="channel id of the first training channel (RNG stream)")
    ap.add_argument("--batch", type=int, default=512, help="training minibatch")
    ap.add_argument("--grid", type=int, default=0, help="simulate every channel on this grid size (device ensembles)")
    ap.add_argument("--save-data", default="", help="write the simulated datasets (npz) and stop if --members 0")
    ap.add_argument("--load-data", default="", help="train and evaluate on datasets written by --save-data")
    args = ap.parse_args()
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(args.seed)
    report = {"args": vars(args)}

    def simulate(channels, realizations, seed, first_id):
        """Chunked simulation with a progress line per chunk; returns dataset arrays.
        With --grid, every channel runs on that grid size through the device
        ensembles (sdata.simulate_ensembles)."""
        xs, ts, vs = [], [], []
        for c0 in range(0, len(channels), args.chunk):
            part = channels[c0:c0 + args.chunk]
            ids = range(first_id + c0, first_id + c0 + len(part))
            if args.grid:
                sim = sdata.simulate_ensembles(part, realizations, seed=seed, device=dev, channel_ids=ids,
                                               grid=args.grid)
                x, t, tv = sdata.dataset_from_hist(part, sim)
            else:
                sim = sdata.simulate(part, realizations, seed=seed, device=dev, channel_ids=ids)
                x, t, tv = sdata.dataset(part, sim)
            xs.append(x), ts.append(t), vs.append(tv)
            print(f"  simulated channels {c0 + len(part)}/{len(channels)} x {realizations}", flush=True)
        return np.concatenate(xs), np.concatenate(ts), np.concatenate(vs)

    if args.load_data:
        # comma-separated files: training pairs concatenated, the test set of
        # the first file that has one
        parts = [np.load(f) for f in args.load_data.split(",")]
        x, t, tv = (np.concatenate([d[k] for d in parts]) for k in ("x", "t", "tv"))
        with_test = [d for d in parts if len(d["xe"])]
        xe, te, tve = (with_test[0][k] for k in ("xe", "te", "tve"))
        metas = [json.loads(str(d["meta"])) for d in parts]
        report["train_simulation_s"] = sum(m["train_simulation_s"] for m in metas)
        report["train_realizations"] = sum(m["train_realizations"] for m in metas)
        report["data_args"] = [m["data_args"] for m in metas]
        t_test = json.loads(str(with_test[0]["meta"]))["test_simulation_s"]
        print(f"loaded {len(x)} training and {len(xe)} test pairs", flush=True)
    else:
        t0 = time.perf_counter()
        train_ch = sdata.sample_channels(args.train_channels, seed=args.seed)
        x, t, tv = simulate(train_ch, args.train_real, args.seed, args.first_id)
        torch.cuda.synchronize()
        t_sim = time.perf_counter() - t0
        report["train_simulation_s"] = t_sim
        report["train_realizations"] = args.train_channels * args.train_real
        print(f"simulated {args.train_channels} training channels x {args.train_real} in {t_sim:.1f} s", flush=True)

        test_ch = sdata.sample_channels(args.test_channels, seed=args.seed + 1)
        t0 = time.perf_counter()
        if args.test_channels:
            xe, te, tve = simulate(test_ch, args.test_real, args.seed + 1, 10 ** 6)
        else:
            xe, te, tve = np.zeros((0, 5), np.float32), np.zeros((0, NB), np.float32), np.zeros(0, np.float32)
        torch.cuda.synchronize()
        t_test = time.perf_counter() - t0
        report["test_simulation_s"] = t_test
        print(f"simulated {args.test_channels} test channels x {args.test_real} in {t_test:.1f} s", flush=True)
        if args.save_data:
            meta = {k: report[k] for k in ("train_simulation_s", "train_realizations", "test_simulation_s")}
            meta["data_args"] = vars(args)
            np.savez_compressed(args.save_data, x=x, t=t, tv=tv, xe=xe, te=te, tve=tve, meta=json.dumps(meta))
            print(f"wrote {args.save_data}", flush=True)
        if args.members == 0:
            return

    members = []
    t0 = time.perf_counter()
    for m in range(args.members):
        net = PDTSurrogate(width=args.width, depth=4, seed=args.seed + 17 * m, device=dev)
        hist = net.train(x, t, tv, steps=args.steps, batch=args.batch, seed=m, log_every=500)
        print(f"member {m}: loss {hist[0]:.3f} -> {hist[-1]:.3f}", flush=True)
        members.append(net)
        if args.out:
            os.makedirs(args.out, exist_ok=True)
            net.save(os.path.join(args.out, f"pdt_member{m}.npz"))
    torch.cuda.synchronize()
    report["train_s"] = time.perf_counter() - t0

    ens = SurrogateEnsemble(members)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p, lv, spread, fallback = ens.pdt(xe)
    torch.cuda.synchronize()
    t_inf = time.perf_counter() - t0
    ks = ks_binned(p, te)
    ks_mean = ks_binned(np.broadcast_to(t.mean(axis=0), te.shape), te)
    # nearest training row in encoded space (same aperture radius grid)
    nearest = np.empty(len(xe), np.int64)
    for i0 in range(0, len(xe), 256):          # blocks keep the distance matrix small
        d2 = ((xe[i0:i0 + 256, None, :] - x[None, :, :]) ** 2).sum(-1)
        nearest[i0:i0 + 256] = np.argmin(d2, axis=1)
    ks_nn = ks_binned(t[nearest], te)
    floor = noise_floor(te, args.test_real, rng)
    report.update({
        "heldout_pairs": int(len(xe)),
        "ks_surrogate_mean": float(ks.mean()), "ks_surrogate_median": float(np.median(ks)),
        "ks_surrogate_p90": float(np.quantile(ks, 0.9)),
        "ks_surrogate_accepted_mean": float(ks[~fallback].mean()) if np.any(~fallback) else None,
        "fallback_fraction": float(fallback.mean()),
        "ks_global_mean_pdt": float(ks_mean.mean()), "ks_nearest_training_channel": float(ks_nn.mean()),
        "ks_sampling_floor": floor,
        "logvar_rmse": float(np.sqrt(np.mean((lv - tve) ** 2))),
        "surrogate_s_per_channel": t_inf / len(xe) * len(sdata.RADII),
        "simulation_s_per_channel": t_test / (len(xe) // len(sdata.RADII)),
    })
    report["speedup_vs_simulation"] = report["simulation_s_per_channel"] / report["surrogate_s_per_channel"]
    print(json.dumps(report, indent=1), flush=True)
    if args.out:
        with open(os.path.join(args.out, "report.json"), "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
