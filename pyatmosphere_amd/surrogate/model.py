"""Transmittance-PDT surrogate: a small MLP trained and evaluated with the
project's own HIP kernels (libpa_hip.so pa_mlp_*).

Inputs (5, encoded to [-1, 1]-ish): log10 Cn2, path length L, beam waist w0,
aperture radius a, wavelength.  Outputs: NB = 64 logits of the transmittance
PDT over [0, 1] (softmax -> bin probabilities) and log of the beam-wander
variance sigma_BW,x^2 = <x_c^2> (simulations/beam.py:35-38).

Training: float32 master weights, float32 GEMMs on the vector units, softmax
cross-entropy + a ranked-probability (squared CDF distance) term + weighted
log-variance MSE, Adam.  Inference: bf16 MFMA GEMMs
with float32 accumulation and fused bias + SiLU.
"""
from __future__ import annotations

from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from .. import _lib
from .._lib import call, ptr

NB = 64
IN_FEATURES = 5
IN_PAD = 16
DOMAIN = {
    "log10_cn2": (-16.0, -14.0),
    "L": (1e3, 50e3),
    "w0": (0.02, 0.2),
    "a": (0.05, 0.3),
    "wvl": (808e-9, 1550e-9),
}


def encode(log10_cn2, L, w0, a, wvl) -> np.ndarray:
    """(M, 5) float32 network inputs scaled to [-1, 1] over the training domain
    (the path length on a log scale)."""
    cols = []
    for key, v in (("log10_cn2", log10_cn2), ("L", L), ("w0", w0), ("a", a), ("wvl", wvl)):
        lo, hi = DOMAIN[key]
        v = np.asarray(v, np.float64)
        if key == "L":
            v, lo, hi = np.log(v), np.log(lo), np.log(hi)
        cols.append((v - lo) / (hi - lo) * 2 - 1)
    return np.stack(cols, axis=-1).astype(np.float32)


def in_domain(x: np.ndarray, tol: float = 1e-6) -> np.ndarray:
    return np.all(np.abs(x) <= 1 + tol, axis=-1)


class PDTSurrogate:
    """MLP with `depth` linear layers of width `width` (SiLU between)."""

    def __init__(self, width: int = 256, depth: int = 4, seed: int = 0, device=None, wv: float = 1.0,
                 wc: float = 4.0):
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.sizes = [IN_PAD] + [width] * (depth - 1) + [NB + 1]
        self.wv, self.wc = wv, wc
        rng = np.random.default_rng(seed)
        self.params: List[Tuple[torch.Tensor, torch.Tensor]] = []
        for i, (fi, fo) in enumerate(zip(self.sizes[:-1], self.sizes[1:])):
            fan_in = IN_FEATURES if i == 0 else fi
            W = rng.normal(0, np.sqrt(2.0 / fan_in), size=(fo, fi)).astype(np.float32)
            if i == 0:
                W[:, IN_FEATURES:] = 0.0
            if i == len(self.sizes) - 2:
                W *= 0.1
            b = np.zeros(fo, np.float32)
            self.params.append((torch.from_numpy(W).to(self.device), torch.from_numpy(b).to(self.device)))
        self.m = [(torch.zeros_like(W), torch.zeros_like(b)) for W, b in self.params]
        self.v = [(torch.zeros_like(W), torch.zeros_like(b)) for W, b in self.params]
        self.t = 0
        self._bf16 = None

    # ------------------------------------------------------------ io
    def state(self) -> dict:
        d = {"sizes": np.array(self.sizes), "wv": np.array(self.wv), "wc": np.array(self.wc)}
        for i, (W, b) in enumerate(self.params):
            d[f"W{i}"] = W.cpu().numpy()
            d[f"b{i}"] = b.cpu().numpy()
        return d

    def save(self, path: str) -> None:
        np.savez(path, **self.state())

    @classmethod
    def load(cls, path: str, device=None) -> "PDTSurrogate":
        z = np.load(path, allow_pickle=False)
        sizes = [int(s) for s in z["sizes"]]
        net = cls(width=sizes[1], depth=len(sizes) - 1, device=device, wv=float(z["wv"]),
                  wc=float(z["wc"]) if "wc" in z else 0.0)
        net.params = [(torch.from_numpy(z[f"W{i}"]).to(net.device), torch.from_numpy(z[f"b{i}"]).to(net.device))
                      for i in range(len(sizes) - 1)]
        net.m = [(torch.zeros_like(W), torch.zeros_like(b)) for W, b in net.params]
        net.v = [(torch.zeros_like(W), torch.zeros_like(b)) for W, b in net.params]
        return net

    def numpy_params(self):
        """float64 (W, b) pairs with the input padding removed (oracle layout)."""
        out = []
        for i, (W, b) in enumerate(self.params):
            Wn = W.cpu().numpy().astype(np.float64)
            if i == 0:
                Wn = Wn[:, :IN_FEATURES]
            out.append((Wn, b.cpu().numpy().astype(np.float64)))
        return out

    # ------------------------------------------------------------ training kernels
    def _pad(self, x: np.ndarray) -> torch.Tensor:
        xp = np.zeros((x.shape[0], IN_PAD), np.float32)
        xp[:, :IN_FEATURES] = x
        return torch.from_numpy(xp).to(self.device)

    def _forward(self, X: torch.Tensor):
        s = _lib.stream_handle()
        M = X.shape[0]
        hs, zs = [X], []
        for i, (W, b) in enumerate(self.params):
            fo, fi = W.shape
            Z = torch.empty(M, fo, dtype=torch.float32, device=self.device)
            call("pa_mlp_gemm_f32", M, fo, fi, ptr(hs[-1]), fi, 0, ptr(W), fi, 1, ptr(Z), fo, 1.0, 0.0, s)
            Y = torch.empty_like(Z)
            call("pa_mlp_bias_act", M, fo, ptr(Z), ptr(b), ptr(Y), int(i < len(self.params) - 1), s)
            zs.append(Z)
            hs.append(Y)
        return hs, zs

    def loss_and_grads(self, X: torch.Tensor, T: torch.Tensor, TV: torch.Tensor):
        """Mean loss (device scalar) and gradients [(dW, db)] for one batch."""
        s = _lib.stream_handle()
        M = X.shape[0]
        hs, zs = self._forward(X)
        out = hs[-1]
        loss = torch.empty(M, dtype=torch.float32, device=self.device)
        d = torch.empty_like(out)
        call("pa_mlp_head_loss", M, NB, ptr(out), ptr(T), ptr(TV), float(self.wv), float(self.wc), ptr(loss),
             ptr(d), s)
        grads = [None] * len(self.params)
        for i in reversed(range(len(self.params))):
            W, b = self.params[i]
            fo, fi = W.shape
            if i < len(self.params) - 1:
                dz = torch.empty_like(d)
                call("pa_mlp_act_backward", M, fo, ptr(zs[i]), ptr(d), ptr(dz), s)
                d = dz
            dW = torch.empty_like(W)
            db = torch.empty_like(b)
            call("pa_mlp_gemm_f32", fo, fi, M, ptr(d), fo, 1, ptr(hs[i]), fi, 0, ptr(dW), fi, 1.0, 0.0, s)
            call("pa_mlp_bias_grad", M, fo, ptr(d), ptr(db), s)
            grads[i] = (dW, db)
            if i > 0:
                dh = torch.empty(M, fi, dtype=torch.float32, device=self.device)
                call("pa_mlp_gemm_f32", M, fi, fo, ptr(d), fo, 0, ptr(W), fi, 0, ptr(dh), fi, 1.0, 0.0, s)
                d = dh
        return loss.mean(), grads

    def adam_step(self, grads, lr: float, b1=0.9, b2=0.999, eps=1e-8) -> None:
        s = _lib.stream_handle()
        self.t += 1
        for (W, b), (gW, gb), (mW, mb), (vW, vb) in zip(self.params, grads, self.m, self.v):
            call("pa_mlp_adam", W.numel(), ptr(W), ptr(gW), ptr(mW), ptr(vW), lr, b1, b2, eps, self.t, s)
            call("pa_mlp_adam", b.numel(), ptr(b), ptr(gb), ptr(mb), ptr(vb), lr, b1, b2, eps, self.t, s)
        self._bf16 = None

    def train(self, x: np.ndarray, target: np.ndarray, tvar: np.ndarray, steps: int, batch: int = 512,
              lr: float = 2e-3, lr_final: float = 5e-5, seed: int = 0, log_every: int = 0) -> List[float]:
        """Adam over shuffled minibatches with a cosine learning-rate decay."""
        rng = np.random.default_rng(seed)
        X = self._pad(x)
        T = torch.from_numpy(np.ascontiguousarray(target, np.float32)).to(self.device)
        TV = torch.from_numpy(np.ascontiguousarray(tvar, np.float32)).to(self.device)
        n = x.shape[0]
        hist = []
        for it in range(steps):
            idx = torch.from_numpy(rng.integers(0, n, size=min(batch, n))).to(self.device)
            loss, grads = self.loss_and_grads(X[idx].contiguous(), T[idx].contiguous(), TV[idx].contiguous())
            frac = it / max(1, steps - 1)
            lr_t = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * frac))
            self.adam_step(grads, lr_t)
            if log_every and (it % log_every == 0 or it == steps - 1):
                hist.append(float(loss.item()))
        return hist

    # ------------------------------------------------------------ inference (bf16 MFMA)
    def _bf16_weights(self):
        if self._bf16 is None:
            s = _lib.stream_handle()
            layers = []
            for i, (W, b) in enumerate(self.params):
                fo, fi = W.shape
                fo_p = ((fo + 31) // 32) * 32
                Wp = torch.zeros(fo_p, fi, dtype=torch.float32, device=self.device)
                Wp[:fo] = W
                bp = torch.zeros(fo_p, dtype=torch.float32, device=self.device)
                bp[:fo] = b
                Wb = torch.empty(fo_p, fi, dtype=torch.int16, device=self.device)
                call("pa_mlp_to_bf16", Wp.numel(), ptr(Wp), ptr(Wb), s)
                layers.append((Wb, bp, fo, fo_p, fi))
            self._bf16 = layers
        return self._bf16

    def predict_raw(self, x: np.ndarray) -> np.ndarray:
        """(M, NB + 1) outputs on bf16 MFMA."""
        s = _lib.stream_handle()
        h = self._pad(x)
        M = h.shape[0]
        layers = self._bf16_weights()
        for i, (Wb, bp, fo, fo_p, fi) in enumerate(layers):
            y = torch.empty(M, fo_p, dtype=torch.float32, device=self.device)
            call("pa_mlp_layer_bf16", M, fo_p, fi, ptr(h), ptr(Wb), ptr(bp), ptr(y), int(i < len(layers) - 1), s)
            h = y
        return h[:, :NB + 1].cpu().numpy().astype(np.float64)

    def predict(self, x: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
        """Bin probabilities (M, NB) of the transmittance PDT and log sigma_BW,x^2."""
        out = self.predict_raw(x)
        lg = out[:, :NB]
        p = np.exp(lg - lg.max(axis=1, keepdims=True))
        return p / p.sum(axis=1, keepdims=True), out[:, NB]


def ks_binned(p: np.ndarray, q: np.ndarray) -> np.ndarray:
    """Kolmogorov-Smirnov distance between binned distributions over the same
    bins (max |CDF difference| at the bin edges)."""
    return np.abs(np.cumsum(p, axis=-1) - np.cumsum(q, axis=-1)).max(axis=-1)


class SurrogateEnsemble:
    """Members trained from different seeds; the spread of their PDTs is the
    confidence signal.  `pdt(...)` returns the ensemble mean, and a mask of the
    rows whose inputs are outside the training domain or whose members disagree
    by more than `max_spread` (KS between a member and the mean), which the
    caller answers with the full simulation."""

    def __init__(self, members: Sequence[PDTSurrogate], max_spread: float = 0.04):
        self.members = list(members)
        self.max_spread = max_spread

    def pdt(self, x: np.ndarray):
        preds = [m.predict(x) for m in self.members]
        ps = np.stack([p for p, _ in preds])
        mean = ps.mean(axis=0)
        spread = np.max([ks_binned(p, mean) for p in ps], axis=0)
        lv = np.mean([v for _, v in preds], axis=0)
        fallback = (spread > self.max_spread) | ~in_domain(x)
        return mean, lv, spread, fallback
