"""Train and evaluate the PDT surrogate on the device.

    python -m pyatmosphere_amd.surrogate.train --train-channels 1500 --train-real 256 \
        --test-channels 64 --test-real 2000 --members 3 --out pyatmosphere_amd/surrogate/weights

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
    args = ap.parse_args()
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(args.seed)
    report = {"args": vars(args)}

    t0 = time.perf_counter()
    train_ch = sdata.sample_channels(args.train_channels, seed=args.seed)
    sim = sdata.simulate(train_ch, args.train_real, seed=args.seed, device=dev)
    torch.cuda.synchronize()
    t_sim = time.perf_counter() - t0
    report["train_simulation_s"] = t_sim
    report["train_realizations"] = args.train_channels * args.train_real
    print(f"simulated {args.train_channels} training channels x {args.train_real} in {t_sim:.1f} s", flush=True)
    x, t, tv = sdata.dataset(train_ch, sim)

    test_ch = sdata.sample_channels(args.test_channels, seed=args.seed + 1)
    t0 = time.perf_counter()
    sim_te = sdata.simulate(test_ch, args.test_real, seed=args.seed + 1, device=dev,
                            channel_ids=range(10 ** 6, 10 ** 6 + args.test_channels))
    torch.cuda.synchronize()
    t_test = time.perf_counter() - t0
    xe, te, tve = sdata.dataset(test_ch, sim_te)
    print(f"simulated {args.test_channels} test channels x {args.test_real} in {t_test:.1f} s", flush=True)

    members = []
    t0 = time.perf_counter()
    for m in range(args.members):
        net = PDTSurrogate(width=args.width, depth=4, seed=args.seed + 17 * m, device=dev)
        hist = net.train(x, t, tv, steps=args.steps, batch=512, seed=m, log_every=500)
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
    d2 = ((xe[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    ks_nn = ks_binned(t[np.argmin(d2, axis=1)], te)
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
        "simulation_s_per_channel": t_test / args.test_channels,
    })
    report["speedup_vs_simulation"] = report["simulation_s_per_channel"] / report["surrogate_s_per_channel"]
    print(json.dumps(report, indent=1), flush=True)
    if args.out:
        with open(os.path.join(args.out, "report.json"), "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
