"""pa_field_moments_static (the entry point that allocates nothing: its
partials live in the library's static device array) against
pa_field_partials + pa_moments_finalize with a caller-allocated partials
buffer, through the C ABI: identical moments, on two calls in a row (the
second with a smaller input), with pupils and with centroid-tracked pupils."""
import numpy as np
import pytest
import torch

from pyatmosphere_amd import _lib
from pyatmosphere_amd._lib import call, ptr

pytestmark = pytest.mark.gpu


def _inputs(n, batch, delta, npup, dev, seed):
    g = torch.Generator().manual_seed(seed)
    x = torch.linspace(-1, 1, n, dtype=torch.float64)
    # shifted Gaussian beams with random phases: non-trivial centroids
    u = torch.empty(batch, n, n, dtype=torch.complex64)
    for b in range(batch):
        cx, cy = (torch.rand(2, generator=g, dtype=torch.float64) - 0.5) * 0.6
        amp = torch.exp(-((x[None, :] - cx) ** 2 + (x[:, None] - cy) ** 2) / 0.08)
        ph = torch.rand(n, n, generator=g, dtype=torch.float64) * 6.28
        u[b] = (amp * torch.exp(1j * ph)).to(torch.complex64)
    chp = torch.tensor([[delta, 0.0, 0.0, 0.0]], dtype=torch.float64)
    pup = torch.zeros(1, _lib.PA_MAX_PUPILS, 3, dtype=torch.float64)
    for p in range(npup):
        pup[0, p] = torch.tensor([(0.1 + 0.1 * p) * n * delta, 0.01 * p, -0.02 * p])
    return u.to(dev), chp.to(dev), pup.to(dev)


def _with_buffer(n, u, chp, pup, npup, pm, batch, stream):
    T = call("pa_tile_rows", n)
    part = torch.empty(batch, n // T, _lib.PA_PART_SUMS, dtype=torch.float64, device=u.device)
    out = torch.empty(batch, _lib.PA_NUM_MOMENTS, dtype=torch.float64, device=u.device)
    call("pa_field_partials", n, ptr(u), ptr(part), ptr(chp), ptr(pup), npup, ptr(pm), None, batch, stream)
    call("pa_moments_finalize", n, ptr(part), ptr(chp), npup, None, ptr(out), batch, stream)
    return out


@pytest.mark.parametrize("n,b1,b2", [(256, 6, 3), (2048, 5, 2)])
def test_static_moments_equal_buffered(dev, n, b1, b2):
    stream = _lib.stream_handle()
    for batch, seed in ((b1, 1), (b2, 2)):          # second call: smaller input
        npup = 3
        u, chp, pup = _inputs(n, batch, 1.5e-3, npup, dev, seed)
        out = torch.full((batch, _lib.PA_NUM_MOMENTS), np.nan, dtype=torch.float64, device=dev)
        print(f"static moments n={n} batch={batch}", flush=True)
        call("pa_field_moments_static", n, ptr(u), ptr(chp), ptr(pup), npup, None, None, ptr(out), batch, stream)
        ref = _with_buffer(n, u, chp, pup, npup, None, batch, stream)
        torch.cuda.synchronize()
        assert torch.equal(out[:, :7 + npup], ref[:, :7 + npup])
        assert torch.all(out[:, 0] > 0)
        # centroid-tracked pupils: the moments just computed shift each field's pupils
        out2 = torch.empty_like(out)
        call("pa_field_moments_static", n, ptr(u), ptr(chp), ptr(pup), npup, ptr(out), None, ptr(out2), batch, stream)
        ref2 = _with_buffer(n, u, chp, pup, npup, ref, batch, stream)
        torch.cuda.synchronize()
        assert torch.equal(out2[:, :7 + npup], ref2[:, :7 + npup])


def test_static_moments_reject_oversized_batch(dev):
    n = 4096
    cap = call("pa_field_moments_static_max", n)
    assert cap == 36
    with pytest.raises(RuntimeError, match="static scratch"):
        call("pa_field_moments_static", n, 1, 1, None, 0, None, None, 1, cap + 1, _lib.stream_handle())
