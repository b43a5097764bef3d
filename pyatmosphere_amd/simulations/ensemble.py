"""Per-channel Monte Carlo ensembles of many channels at once.

`run_mixed` advances every slot of a mixed-channel batch (one channel per
slot, any grid sizes and screen counts) by one realization per replay of one
captured graph, and accumulates on the device, per channel, the BeamResult
quantities (float64 sums of mean_x^2, 4 mean_x2, ... and their squares,
simulations/beam.py:35-71) and each pupil's 64-bin transmittance histogram
(simulations/pdt.py:28-48), until every channel has `beam_samples` beam
samples and `pdt_samples` PDT samples.  Nothing per realization comes back to
the host.

    python -m pyatmosphere_amd.simulations.ensemble --beam 2000 --pdt 6000 --out ensembles.json

runs the 256-slot mixed workload (512^2 / 1024^2 / 2048^2 grids, 3-10
screens) to completion and writes the per-channel statistics.
"""
import argparse
import json
import sys
import time
from typing import Dict, Optional, Sequence

import numpy as np
import torch

from .. import engine


def run_mixed(specs: Sequence[engine.ChannelSpec], beam_samples: int, pdt_samples: int, seed: int = 0,
              mode: str = "f16x3", device: Optional[torch.device] = None, progress=None) -> Dict:
    """Run one realization per channel per replay until every channel has its
    samples; returns per-channel beam statistics, PDT counts and densities,
    and the timing."""
    dev = device or torch.device("cuda", torch.cuda.current_device())
    mb = engine.MixedBatch(specs, np.arange(len(specs)), seed=seed, mode=mode, device=dev)
    ens = mb.ensemble(beam_samples, pdt_samples)
    S = len(specs)
    replays = max(beam_samples, pdt_samples)
    mb.run(np.arange(S))                          # capture (its moments are not kept)
    ens.counts.zero_()
    ens.acc.zero_()
    ens.hist.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in range(replays):
        mb.run(np.arange(S) + (r + 1) * S, ensemble=ens)
        if progress is not None and (r + 1) % 500 == 0:
            torch.cuda.synchronize()
            progress(r + 1, replays, time.perf_counter() - t0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    if not ens.complete():
        raise RuntimeError("ensemble incomplete after the planned replays")
    dens, edges = ens.pdt()
    return {"beam": ens.beam(), "pdt_counts": ens.histogram(), "pdt_density": dens, "edges": edges,
            "realizations": replays * S, "seconds": dt, "realizations_per_s": replays * S / dt}


def main(argv=None):
    from .. import workloads
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--beam", type=int, default=2000)
    ap.add_argument("--pdt", type=int, default=6000)
    ap.add_argument("--slots", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)
    specs = workloads.mixed_channels(args.slots)

    def progress(r, total, el):
        print(f"replay {r}/{total}: {el:.1f} s", flush=True)

    res = run_mixed(specs, args.beam, args.pdt, seed=args.seed, progress=progress)
    summary = {
        "slots": len(specs), "beam_samples": args.beam, "pdt_samples": args.pdt,
        "realizations": res["realizations"], "seconds": res["seconds"],
        "realizations_per_s": res["realizations_per_s"],
        "channels": [{"n": s.n, "delta": s.delta, "screens": s.nscreens, "length": s.length, "wvl": s.wvl,
                      "w0": s.w0, "pupil": s.pupils[0][0], "bw": b["bw"], "lt": b["lt"], "st": b["st"],
                      "pdt_counts": res["pdt_counts"][i, 0].tolist()}
                     for i, (s, b) in enumerate(zip(specs, res["beam"]))],
    }
    print(json.dumps({k: v for k, v in summary.items() if k != "channels"}), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(summary, f)
    return 0


if __name__ == "__main__":
    sys.exit(main())
