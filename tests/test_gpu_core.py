"""Device parity of the split-step kernels against the FP64 oracle."""
import dataclasses

import numpy as np
import pytest
import torch

from oracle import fp64 as o
from oracle.philox import spectral_draw
from pyatmosphere_amd import engine
from tests.helpers import channel_spec, host_spectra, oracle_field, rel_l2

pytestmark = pytest.mark.gpu


def _vacuum_spec(n, delta, wvl, L):
    return engine.ChannelSpec(n=n, delta=delta, wvl=wvl, w0=1.0, F0=np.inf, length=L, positions=np.zeros(0),
                              points=16, base=np.ones(16, np.float32), psd=np.zeros(16, np.float32))


@pytest.mark.parametrize("n", [64, 128, 256, 512, 1024, 2048, 4096])
def test_vacuum_step_parity(dev, n):
    """row_first + col_step + row_last == ifft2(H fft2 u) of the oracle (fp32 FFTs)."""
    print(f"vacuum step n={n}", flush=True)
    rng = np.random.default_rng(n)
    u = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))).astype(np.complex64)
    delta, wvl, L = 0.002, 1.55e-6, 3e3
    spec = _vacuum_spec(n, delta, wvl, L)
    tables = engine.DeviceTables([spec], dev)
    ws = engine.Workspace(n, 16, 1, "f16x3", dev)
    out = torch.empty(1, n, n, dtype=torch.complex64, device=dev)
    engine.run_batch(ws, tables, np.zeros(1), np.zeros(1), field_in=torch.from_numpy(u).to(dev)[None],
                     field_out=out, reduce=False)
    ref = o.vacuum_step(u.astype(np.complex128), delta, wvl, L)
    err = rel_l2(out[0].cpu().numpy(), ref)
    print(f"  rel-L2 {err:.2e}", flush=True)
    assert err < 2e-6


@pytest.mark.parametrize("n", [64, 128, 256, 512, 1024, 2048, 4096])
def test_source_fresnel_matches_oracle(dev, n):
    """pa_source_fresnel: the outer product of its two vectors is the FP64
    vacuum step of the Gaussian source over the first step length, the field at
    the first screen plane (theory/vacuum.py:5-7 on sources.py:21-23), for a
    collimated and a focused beam and two channels of one batch."""
    delta = 0.002 if n <= 512 else 0.0008
    specs = [channel_spec(n, delta, 808e-9, n * delta / 10, 2e3, 3, 64, 1e-15),
             channel_spec(n, delta * 0.75, 1550e-9, n * delta / 12, 5e3, 4, 64, 1e-15, F0=1.5e3)]
    tables = engine.DeviceTables(specs, dev)
    ws = engine.Workspace(n, 64, 2, "f16x3", dev)
    run = engine.StepRunner(ws, tables, np.array([0, 1]), np.array([3, 4]))
    run.begin(None, separable=True)
    assert run.sep0
    torch.cuda.synchronize()
    sep = ws.sep[:2].cpu().numpy()
    for b, sp in enumerate(specs):
        z1 = o.step_lengths(sp.length, sp.positions)[0]
        ref = o.vacuum_step(o.gaussian_source(n, sp.delta, sp.wvl, sp.w0, sp.F0), sp.delta, sp.wvl, z1)
        err = rel_l2(np.outer(sep[b, 0], sep[b, 1]), ref)
        print(f"source_fresnel n={n} channel {b}: rel-L2 {err:.2e}", flush=True)
        assert err < 2e-6


@pytest.mark.parametrize("n", [256, 2048])
def test_separable_start_matches_2d_start(dev, n):
    """run_batch with the separable first step (pa_source_fresnel +
    pa_row_mid_source) against the 2-D first step (pa_row_first + pa_col_step
    + pa_row_mid) on the same draws: the final fields agree to float32
    rounding and the moments to 1e-5, over two channels with different screen
    counts in one batch; the third has path losses and its first screen at
    z = 0 (a zero-length first step that only carries the loss)."""
    delta = 0.002 if n <= 512 else 0.001
    lossy = channel_spec(n, delta, 808e-9, n * delta / 9, 3e3, 3, 256, 1e-15, position="before",
                         pupils=[(n * delta / 7, 0, 0)])
    lossy = dataclasses.replace(lossy, gains=[0.8, 0.9, 0.95, 0.7])
    specs = [channel_spec(n, delta, 808e-9, n * delta / 10, 4e3, 4, 256, 2e-15, pupils=[(n * delta / 8, 0, 0)]),
             channel_spec(n, delta, 1550e-9, n * delta / 12, 6e3, 2, 256, 1e-15, F0=4e3,
                          pupils=[(n * delta / 6, 0, 0)]), lossy]
    assert lossy.step_lengths()[0] == 0.0
    tables = engine.DeviceTables(specs, dev)
    ws = engine.Workspace(n, 256, 4, "f16x3", dev)
    chan = np.array([0, 1, 2, 0])
    ns = np.array([specs[c].nscreens for c in chan])
    res = []
    for sep in (True, False):
        out = torch.empty(4, n, n, dtype=torch.complex64, device=dev)
        m = engine.run_batch(ws, tables, chan, ns, seed=7, rid=np.arange(4), field_out=out, separable=sep)
        torch.cuda.synchronize()
        res.append((out.cpu().numpy(), m.cpu().numpy()))
    for b in range(4):
        err = rel_l2(res[0][0][b], res[1][0][b])
        print(f"separable vs 2-D start n={n} element {b}: field rel-L2 {err:.2e}", flush=True)
        assert err < 2e-6
    assert np.allclose(res[0][1][:, :8], res[1][1][:, :8], rtol=1e-5, atol=1e-12)


def test_fft_identity_step(dev):
    """A zero-length step is the identity (pathes.py:30, 39-40): the kernels run
    forward and inverse transforms with H = 1/N^2 and must return the input."""
    n = 512
    rng = np.random.default_rng(1)
    u = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))).astype(np.complex64)
    spec = _vacuum_spec(n, 0.001, 1e-6, 0.0)
    tables = engine.DeviceTables([spec], dev)
    ws = engine.Workspace(n, 16, 1, "f16x3", dev)
    out = torch.empty(1, n, n, dtype=torch.complex64, device=dev)
    engine.run_batch(ws, tables, np.zeros(1), np.zeros(1), field_in=torch.from_numpy(u).to(dev)[None],
                     field_out=out, reduce=False)
    assert rel_l2(out[0].cpu().numpy(), u) < 1e-6


def test_philox_draw_bitwise(dev):
    P = 1024
    spec = channel_spec(256, 0.002, 808e-9, 0.05, 10e3, 5, P, 1e-15)
    tables = engine.DeviceTables([spec], dev)
    ws = engine.Workspace(256, P, 3, "f16x3", dev)
    rid = np.array([0, 17, 2 ** 40 + 5])
    ws.chan[:3].zero_()
    ws.rid[:3].copy_(torch.from_numpy(rid))
    engine.draw_device(ws, tables, 987654321, 2, 3, None)
    sq = np.sqrt(spec.psd)
    for b in range(3):
        rho, theta, value = spectral_draw(int(rid[b]), 2, 987654321, spec.base, sq)
        assert np.array_equal(ws.theta[b].cpu().numpy(), theta)
        assert np.allclose(ws.rho[b].cpu().numpy(), rho, rtol=1.2e-7, atol=0)
        assert np.allclose(ws.value[b].cpu().numpy(), value, rtol=1e-6, atol=1e-30)


@pytest.mark.parametrize("P,shift", [(1024, (0.0, 0.0)), (1500, (0.37, -1.9))])
def test_fused_draw_prep_matches_draw_then_prep(dev, P, shift):
    """pa_screen_draw_prep gives the fxy and sparam of pa_spectral_draw followed
    by pa_screen_prep bit for bit (P = 1500: a partial second chunk of 1024
    points; a shifted channel rotates the amplitudes), and leaves elements whose
    screen count is reached untouched."""
    spec = dataclasses.replace(channel_spec(256, 0.002, 808e-9, 0.05, 10e3, 5, P, 1e-15), shift=shift)
    tables = engine.DeviceTables([spec], dev)
    B = 3
    ws = engine.Workspace(256, P, B, "f16x3", dev)
    ws.chan[:B].zero_()
    ws.rid[:B].copy_(torch.tensor([0, 17, 2 ** 40 + 5]))
    st = torch.cuda.current_stream().cuda_stream
    engine.draw_device(ws, tables, 987654321, 2, B, st)
    engine.prep_screen(ws, tables, 2, B, None, st)
    fxy, sp = ws.fxy.clone(), ws.sparam[:, :7].clone()
    ws.fxy.fill_(7.0)
    ws.sparam.fill_(7.0)
    nsteps = torch.tensor([5, 2, 5], dtype=torch.int32, device=dev)   # element 1 has no screen 2
    engine.draw_prep_device(ws, tables, 987654321, 2, B, nsteps, st)
    torch.cuda.synchronize()
    for b in (0, 2):
        assert torch.equal(ws.fxy[b], fxy[b]) and torch.equal(ws.sparam[b, :7], sp[b])
    assert bool((ws.fxy[1] == 7.0).all()) and bool((ws.sparam[1] == 7.0).all())


@pytest.mark.parametrize("shift", [(0.0, 0.0), (0.37, -1.9)])
def test_polynomial_block_coefficients(dev, shift):
    """The polynomial block on the device against float64 numpy with scipy's
    Bessel functions, from the device's own fxy (shift-rotated amplitudes), on
    the 2048^2 / 1.5 mm channel: pa_screen_tables grouped the points stably
    (the same multiset; slots [0, n2) turn by at most 10 rad over the grid on
    both axes, [n2, d0) on the column axis only, d0 a multiple of 16); G (fp64
    MFMA over 64-point chunks added in chunk order) and the h Chebyshev
    coefficients from the first n2 slots, the g coefficients from all d0; the
    A scale lowered for the block's column bounds."""
    from scipy.special import jv
    n, P, B, T = 2048, 1024, 2, 32
    spec = dataclasses.replace(channel_spec(n, 0.0015, 808e-9, 0.12, 50e3, 5, P, 5e-16), shift=shift)
    tables = engine.DeviceTables([spec], dev)
    ws = engine.Workspace(n, P, B, "f16x3", dev)
    ws.chan[:B].zero_()
    ws.rid[:B].copy_(torch.arange(B))
    st = torch.cuda.current_stream().cuda_stream
    engine.draw_prep_device(ws, tables, 31, 0, B, None, st)
    torch.cuda.synchronize()
    drawn = ws.fxy.cpu().numpy().copy()
    engine.plane_wave_tables(ws, tables, 0, B, None, st)
    torch.cuda.synchronize()
    fxy, sp = ws.fxy.cpu().numpy(), ws.sparam.cpu().numpy()
    X = n / 2 * spec.delta
    m = np.arange(T)

    def coef(w):
        c = 2 * (1j) ** m[None, :] * jv(m[None, :], (2 * np.pi * w * X)[:, None])
        c[:, 0] = c[:, 0] / 2 - 1
        return c
    for b in range(B):
        d0, n2 = int(sp[b, 5]), int(sp[b, 7])
        f, g0 = fxy[b].reshape(P, 4), drawn[b].reshape(P, 4)
        # a stable grouping of the same points
        assert np.array_equal(np.sort(f.view("V32"), axis=0), np.sort(g0.view("V32"), axis=0))
        col = np.abs(2 * np.pi * f[:, 0]) * X <= 14
        row = np.abs(2 * np.pi * f[:, 1]) * X <= 14
        assert d0 % 16 == 0 and d0 >= 640 and sp[b, 4] == 2 * d0 - 32 and 0 <= d0 - n2 <= 64
        assert np.all(col[:n2] & row[:n2]) and np.all(col[n2:d0] & ~row[n2:d0])
        print(f"shift {shift}: d0 {d0}, both-axis points {n2}, column-axis points {d0 - n2}", flush=True)
        a = f[:, 2] + 1j * f[:, 3]
        Q, Pm = coef(f[:d0, 1]), coef(f[:d0, 0])
        scale = 4 * np.abs(a[:d0]).sum()   # bound on every term sum
        G = np.einsum("k,kn,km->nm", a[:n2], Q[:n2], Pm[:n2]).real
        got = sp[b, 8:8 + T * T].reshape(T, T)
        assert np.abs(got - G).max() < 1e-13 * scale
        assert np.abs(sp[b, 8 + T * T:8 + T * T + T] - (a[:n2] @ Q[:n2]).real).max() < 1e-13 * scale
        assert np.abs(sp[b, 8 + T * T + T:8 + T * T + 2 * T] - (a[:d0] @ Pm).real).max() < 1e-13 * scale
        bound = np.abs(got).sum(axis=0) + 2 * (np.abs(a[n2:d0])[:, None] * np.abs(Pm[n2:])).sum(axis=0)
        expect = min(sp[b, 6], 2.0 ** (14 - np.ceil(np.log2(bound.max()))))
        assert sp[b, 1] == expect and sp[b, 3] == 1.0 / (expect * sp[b, 2])


@pytest.mark.parametrize("mode,tol", [("f16x3", 2e-5), ("f32", 2e-5), ("bf16x3", 2e-2)])
def test_screen_synthesis_parity(dev, mode, tol):
    """Composed screen phi0 + h + g + phi2 against the complex128 plane-wave sum."""
    n, P = 512, 1024
    spec = channel_spec(n, 0.0015, 808e-9, 0.09, 10e3, 5, P, 1e-15)
    spectra, per_real = host_spectra(spec, 1, seed=0)
    tables = engine.DeviceTables([spec], dev)
    ws = engine.Workspace(n, P, 1, mode, dev)
    engine.upload_spectra(ws, *spectra[0], batch=1)
    ws.chan[:1].zero_()
    fac_buf = torch.empty(1, n, n, dtype=torch.complex64, device=dev)
    engine.synth_screen(ws, tables, 0, 1, None, None, phi2=ws.phi2_buffer(), fac=fac_buf)
    phi = torch.empty(1, n, n, dtype=torch.float32, device=dev)
    engine.call("pa_screen_compose", n, engine.ptr(ws.sparam), engine.ptr(ws.hg), engine.ptr(ws.phi2),
                engine.ptr(phi), 1, None)
    ref = o.screen_phase(n, spec.delta, per_real[0][0])
    got = phi[0].cpu().numpy().astype(np.float64)
    # the float32 composed value carries its own rounding of |phi| ~ 1e3 rad
    excess = np.abs(got - ref) - np.abs(ref) * 2 ** -24
    print(f"{mode}: max |phi| {np.abs(ref).max():.1f}, max excess error {excess.max():.2e}", flush=True)
    assert excess.max() < tol
    # the GEMM's optional complex factor output is e^{-i phi} (work layout)
    T = n // ws.tiles   # element (r, c) sits at ((r // T) n + c) T + r % T
    fac = fac_buf[0].cpu().numpy().reshape(ws.tiles, n, T).transpose(0, 2, 1).reshape(n, n)
    assert np.allclose(np.abs(fac), 1.0, atol=2e-6)
    dphase = np.angle(fac * np.exp(1j * ref))
    print(f"{mode}: max factor phase error {np.abs(dphase).max():.2e}", flush=True)
    assert np.abs(dphase).max() < tol + np.abs(ref).max() * 2 ** -23
    # the screen angle the row pass applies: e^{-i ang} = e^{-i phi}, with ang =
    # phi2 + the row and column angles in [-pi, pi] rounded to float32
    ang = ws.ang[0].cpu().numpy().reshape(ws.tiles, n, T).transpose(0, 2, 1).reshape(n, n).astype(np.float64)
    dang = np.angle(np.exp(-1j * ang) * np.exp(1j * ref))
    print(f"{mode}: max |phi2| {np.abs(ang).max():.1f}, max screen-angle phase error {np.abs(dang).max():.2e}",
          flush=True)
    assert np.abs(ang).max() < 64 and np.abs(dang).max() < tol + 4e-6 + np.abs(ref).max() * 2 ** -23
    # the float64 parts alone: phi0 + h + g are exact to float64
    hg = ws.hg[0].cpu().numpy()
    phi0 = ws.sparam[0, 0].item()
    assert abs(phi0 - per_real[0][0].value.real.astype(np.float64).sum()) < 1e-9
    # the low-frequency points went through the polynomial block
    kbeg, d0, n2 = ws.sparam[0, 4].item(), ws.sparam[0, 5].item(), ws.sparam[0, 7].item()
    print(f"{mode}: polynomial block d0 {d0:.0f} ({n2:.0f} on both axes), kbeg {kbeg:.0f} of {2 * P}", flush=True)
    assert d0 >= 32 and d0 % 16 == 0 and kbeg == 2 * d0 - 32 and n2 <= d0


@pytest.mark.parametrize("shift", [(0.37, -1.9), (-12.5, 0.004)])
def test_screen_synthesis_shifted(dev, shift):
    """A shifted screen (phase_screens.py:112-113: x + shift[0], y + shift[1];
    the wind path's shifts), composed on the device against the complex128
    plane-wave sum on the shifted grid.  The polynomial block expands about the
    shifted centre, so this covers its phase factors e^{i w shift}."""
    import dataclasses
    n, P = 512, 1024
    spec = dataclasses.replace(channel_spec(n, 0.0015, 808e-9, 0.09, 10e3, 5, P, 1e-15), shift=shift)
    spectra, per_real = host_spectra(spec, 1, seed=2)
    tables = engine.DeviceTables([spec], dev)
    ws = engine.Workspace(n, P, 1, "f16x3", dev)
    engine.upload_spectra(ws, *spectra[0], batch=1)
    ws.chan[:1].zero_()
    engine.synth_screen(ws, tables, 0, 1, None, None, phi2=ws.phi2_buffer())
    phi = torch.empty(1, n, n, dtype=torch.float32, device=dev)
    engine.call("pa_screen_compose", n, engine.ptr(ws.sparam), engine.ptr(ws.hg), engine.ptr(ws.phi2),
                engine.ptr(phi), 1, None)
    ref = o.screen_phase(n, spec.delta, per_real[0][0], shift=shift)
    excess = np.abs(phi[0].cpu().numpy().astype(np.float64) - ref) - np.abs(ref) * 2 ** -24
    klow = ws.sparam[0, 5].item()
    print(f"shift {shift}: d0 {klow:.0f}, max |phi| {np.abs(ref).max():.1f}, max excess {excess.max():.2e}",
          flush=True)
    assert klow >= 32 and excess.max() < 2e-5


def test_row_pass_exponential(dev):
    """e^{-ix} as the row pass forms the screen factor (pa_expmi: float32
    reduction by 2 pi, hardware sine and cosine) against float64 numpy, over the
    range of the 2-D screen remainder and beyond: |error| < 5e-7."""
    x = np.concatenate([np.linspace(-512, 512, 1 << 21), np.linspace(-np.pi, np.pi, 1 << 16),
                        [0.0, np.pi, -np.pi, 2 * np.pi, 1e-30, -1e-7]]).astype(np.float32)
    xt = torch.from_numpy(x).to(dev)
    out = torch.empty(len(x), dtype=torch.complex64, device=dev)
    engine.call("pa_expmi", len(x), engine.ptr(xt), engine.ptr(out), None)
    got = out.cpu().numpy().astype(np.complex128)
    err = np.abs(got - np.exp(-1j * x.astype(np.float64)))
    print(f"row-pass exponential: max |error| {err.max():.2e}, rms {np.sqrt((err ** 2).mean()):.2e}", flush=True)
    assert err.max() < 5e-7


@pytest.mark.parametrize("P", [16, 48])
def test_screen_small_point_counts(dev, P):
    """P = 16: fewer than 32 qualifying points, every point tabulated directly
    (d0 = 0, kbeg = 0); P = 48: the polynomial block next to few direct points."""
    n = 256
    spec = channel_spec(n, 0.002, 808e-9, 0.05, 5e3, 3, P, 1e-15)
    spectra, per_real = host_spectra(spec, 1, seed=3)
    tables = engine.DeviceTables([spec], dev)
    ws = engine.Workspace(n, P, 1, "f16x3", dev)
    engine.upload_spectra(ws, *spectra[0], batch=1)
    ws.chan[:1].zero_()
    engine.synth_screen(ws, tables, 0, 1, None, None, phi2=ws.phi2_buffer())
    phi = torch.empty(1, n, n, dtype=torch.float32, device=dev)
    engine.call("pa_screen_compose", n, engine.ptr(ws.sparam), engine.ptr(ws.hg), engine.ptr(ws.phi2),
                engine.ptr(phi), 1, None)
    ref = o.screen_phase(n, spec.delta, per_real[0][0])
    excess = np.abs(phi[0].cpu().numpy().astype(np.float64) - ref) - np.abs(ref) * 2 ** -24
    klow = ws.sparam[0, 5].item()
    print(f"P={P}: d0 {klow:.0f}, max excess error {excess.max():.2e}", flush=True)
    assert excess.max() < 2e-5
    if P == 16:
        assert klow == 0 and ws.sparam[0, 4].item() == 0


@pytest.mark.parametrize("cfg", [
    dict(n=256, delta=0.002, wvl=808e-9, w0=0.05, L=5e3, S=3, P=64, Cn2=1e-15),
    dict(n=512, delta=0.0015, wvl=808e-9, w0=0.09, L=10e3, S=5, P=1024, Cn2=1e-15),   # the 512^2 CPU channel
])
def test_channel_field_parity(dev, cfg):
    """Full split-step realization, field-by-field against the FP64 oracle fed the
    same seeded coefficients: relative L2 <= 1e-4."""
    spec = channel_spec(**cfg)
    spectra, per_real = host_spectra(spec, 2, seed=0)
    tables = engine.DeviceTables([spec], dev)
    ws = engine.Workspace(spec.n, spec.points, 2, "f16x3", dev)
    out = torch.empty(2, spec.n, spec.n, dtype=torch.complex64, device=dev)
    mom = engine.run_batch(ws, tables, np.zeros(2), np.full(2, spec.nscreens), host_spectra=spectra,
                           field_out=out, reduce=True).cpu().numpy()
    for b in range(2):
        ref = oracle_field(spec, per_real[b])
        err = rel_l2(out[b].cpu().numpy(), ref)
        print(f"n={spec.n} realization {b}: field rel-L2 {err:.2e}", flush=True)
        assert err < 1e-4
        m = o.moments(ref, spec.delta)
        for i, key in enumerate(engine.MOMENT_NAMES[:6]):
            assert abs(mom[b, i] - m[key]) <= 1e-4 * max(abs(m[key]), m["mean_x2"]), key


def test_channel_field_parity_2048(dev):
    """The 2048^2 / 50 km / 5-screen channel (r = 0.2 m pupil): 16 realizations
    drawn in reference order from np.random seed 0, one batch on the device,
    each field against the FP64 oracle fed the same coefficients: relative
    L2 <= 1e-4, and the pupil transmittance to the same relative bound."""
    R = 16
    spec = channel_spec(n=2048, delta=0.0015, wvl=808e-9, w0=0.12, L=50e3, S=5, P=1024, Cn2=5e-16,
                        pupils=[(0.2, 0.0, 0.0)])
    spectra, per_real = host_spectra(spec, R, seed=0)
    tables = engine.DeviceTables([spec], dev)
    ws = engine.Workspace(2048, 1024, R, "f16x3", dev)
    out = torch.empty(R, 2048, 2048, dtype=torch.complex64, device=dev)
    mom = engine.run_batch(ws, tables, np.zeros(R), np.full(R, 5), host_spectra=spectra, field_out=out,
                           reduce=True).cpu().numpy()
    errs = []
    for b in range(R):
        ref = oracle_field(spec, per_real[b])
        errs.append(rel_l2(out[b].cpu().numpy(), ref))
        eta_ref = o.pupil_eta(ref, spec.delta, 0.2)
        assert abs(mom[b, 7] - eta_ref) <= 1e-4 * max(eta_ref, 1e-3), b
        print(f"2048 realization {b}: field rel-L2 {errs[-1]:.2e}, eta {mom[b, 7]:.6f} (oracle {eta_ref:.6f})",
              flush=True)
    print(f"2048 field rel-L2 over {R} realizations: max {max(errs):.2e}, mean {np.mean(errs):.2e}", flush=True)
    assert max(errs) < 1e-4


def test_moments_and_pupils(dev):
    n, delta = 512, 0.002
    rng = np.random.default_rng(4)
    x, y = o.grid_xy(n, delta)
    u = (np.exp(-((x - 0.05) ** 2 + (y + 0.02) ** 2) / 0.04 ** 2) * (1 + 0.1 * rng.normal(size=(n, n)))
         ).astype(np.complex64)
    # the last two: edge through grid points (12 cells around the grid point
    # (10, -4) cells off centre) and a half-cell radius
    pupils = [(0.03, 0.0, 0.0), (0.05, 0.02, -0.01), (0.2, 0.0, 0.0), (12 * delta, 10 * delta, 4 * delta),
              (delta / 2, 0.0, 0.0)]
    got = engine.field_moments(torch.from_numpy(u).to(dev), delta, pupils)[0]
    m = o.moments(u.astype(np.complex128), delta)
    for i, key in enumerate(engine.MOMENT_NAMES):
        assert np.isclose(got[i], m[key], rtol=1e-9, atol=1e-14), key
    for j, (a, sx, sy) in enumerate(pupils):
        assert np.isclose(got[7 + j], o.pupil_eta(u.astype(np.complex128), delta, a, (sx, sy)), rtol=1e-9)
    tracked = engine.field_moments(torch.from_numpy(u).to(dev), delta, [(0.03, 0, 0)], tracked=True)[0]
    assert np.isclose(tracked[7], o.pupil_eta(u.astype(np.complex128), delta, 0.03,
                                              (m["mean_x"], m["mean_y"])), rtol=1e-9)


def test_fused_reduction_matches_field_reduction(dev):
    # the last three pupils have edges through grid points (radius a whole number
    # of cells, centre on a grid point) and a half-cell radius
    pupils = [(0.05, 0, 0), (0.1, 0.01, 0.02), (25 * 0.002, 0.0, 0.0), (40 * 0.002, 3 * 0.002, -5 * 0.002),
              (0.001, 0.0, 0.0)]
    spec = channel_spec(n=512, delta=0.002, wvl=1550e-9, w0=0.05, L=8e3, S=4, P=256, Cn2=2e-15, pupils=pupils)
    tables = engine.DeviceTables([spec], dev)
    ws = engine.Workspace(512, 256, 4, "f16x3", dev)
    out = torch.empty(4, 512, 512, dtype=torch.complex64, device=dev)
    mom = engine.run_batch(ws, tables, np.zeros(4), np.full(4, 4), seed=11, rid=np.arange(4),
                           field_out=out).cpu().numpy()
    ref = engine.field_moments(out, spec.delta, spec.pupils)
    k = 7 + len(pupils)
    assert np.allclose(mom[:, :k], ref[:, :k], rtol=1e-9, atol=1e-15)


def test_mixed_batch(dev):
    """Elements of different channels and screen counts in one launch equal their
    single-channel runs (same device spectra)."""
    specs = [channel_spec(n=256, delta=0.002, wvl=808e-9, w0=0.05, L=5e3, S=3, P=128, Cn2=1e-15),
             channel_spec(n=256, delta=0.001, wvl=1550e-9, w0=0.03, L=2e3, S=6, P=128, Cn2=5e-15),
             channel_spec(n=256, delta=0.0015, wvl=808e-9, w0=0.08, L=9e3, S=4, P=128, Cn2=2e-16)]
    tables = engine.DeviceTables(specs, dev)
    ws = engine.Workspace(256, 128, 6, "f16x3", dev)
    chan = np.array([0, 1, 2, 1, 0, 2])
    ns = np.array([specs[c].nscreens for c in chan])
    out = torch.empty(6, 256, 256, dtype=torch.complex64, device=dev)
    engine.run_batch(ws, tables, chan, ns, seed=5, rid=np.arange(6), field_out=out, reduce=False)
    for b in range(6):
        t1 = engine.DeviceTables([specs[chan[b]]], dev)
        w1 = engine.Workspace(256, 128, 1, "f16x3", dev)
        o1 = torch.empty(1, 256, 256, dtype=torch.complex64, device=dev)
        engine.run_batch(w1, t1, np.zeros(1), ns[b:b + 1], seed=5, rid=np.array([b]), field_out=o1, reduce=False)
        assert torch.equal(out[b], o1[0]), b


def test_graph_replay_matches_eager(dev):
    """The captured launch sequence replayed for new realization ids equals the
    eager run, replay after replay, with eager work and syncs in between."""
    spec = channel_spec(n=512, delta=0.002, wvl=1550e-9, w0=0.05, L=8e3, S=4, P=256, Cn2=2e-15,
                        pupils=[(0.05, 0, 0)])
    tables = engine.DeviceTables([spec], dev)
    ws_g = engine.Workspace(512, 256, 3, "f16x3", dev)
    ws_e = engine.Workspace(512, 256, 3, "f16x3", dev)
    gb = engine.GraphedBatch(ws_g, tables, np.zeros(3), np.full(3, 4), seed=21)
    for k in range(4):
        rid = np.arange(3) + 10 * k
        got = gb.run(rid).cpu().numpy()
        ref = engine.run_batch(ws_e, tables, np.zeros(3), np.full(3, 4), seed=21, rid=rid).cpu().numpy()
        torch.cuda.synchronize()
        assert np.array_equal(got[:, :8], ref[:, :8]), k


def test_overlapped_schedule_matches_single_stream(dev):
    """Screen synthesis on the side stream into per-screen angle buffers gives
    the same fields and moments, bit for bit, as the single-stream schedule."""
    spec = channel_spec(n=512, delta=0.0015, wvl=808e-9, w0=0.09, L=10e3, S=5, P=1024, Cn2=1e-15,
                        pupils=[(0.05, 0, 0)])
    tables = engine.DeviceTables([spec], dev)
    ws = engine.Workspace(512, 1024, 2, "f16x3", dev)
    outs, moms = [], []
    for overlap in (False, True):
        out = torch.empty(2, 512, 512, dtype=torch.complex64, device=dev)
        m = engine.run_batch(ws, tables, np.zeros(2), np.full(2, 5), seed=8, rid=np.arange(2), field_out=out,
                             overlap=overlap).cpu().numpy()
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy())
        moms.append(m)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(moms[0][:, :8], moms[1][:, :8])


def test_mixed_batch_equals_per_class_runs(dev):
    """Slots of three grid sizes with different screen counts, spacings and
    turbulence in one captured graph: every slot's moments are bit-equal to a
    run_batch of its class with the same realization ids, on two replays."""
    specs = [channel_spec(64, 0.004, 808e-9, 0.05, 2e3, 3, 64, 1e-15),
             channel_spec(128, 0.002, 1550e-9, 0.04, 5e3, 5, 64, 5e-15),
             channel_spec(256, 0.0015, 808e-9, 0.05, 3e3, 4, 128, 2e-15),
             channel_spec(128, 0.003, 808e-9, 0.08, 8e3, 2, 64, 1e-15)]
    slot_channel = [0, 1, 2, 3, 1, 2, 0, 3, 2]
    mb = engine.MixedBatch(specs, slot_channel, seed=5, mode="f16x3", device=dev)
    for rep in range(2):
        rid = np.arange(len(slot_channel)) + 1000 * rep
        got = mb.run(rid).cpu().numpy().copy()
        for c in mb.classes:
            idx = c["slots"]
            key = [specs[i] for i in sorted(set(slot_channel[j] for j in idx))]
            tables = engine.DeviceTables(key, dev)
            ws = engine.Workspace(c["n"], c["points"], len(idx), "f16x3", dev)
            ref = engine.run_batch(ws, tables, c["gb"].chan, c["gb"].nscreens, seed=5, rid=rid[idx],
                                   overlap=False).cpu().numpy()
            assert np.array_equal(got[idx], ref), (rep, c["n"])


@pytest.mark.parametrize("overlap", [False, True])
def test_whole_batch_synthesis_matches_per_screen(dev, monkeypatch, overlap):
    """Synthesizing every screen of a batch at once (the draw per screen, the
    tables and the GEMM over all (screen, realization) pairs) gives the same
    fields and moments, bit for bit, as screen-by-screen synthesis, on a batch
    of two channels with 2 and 5 screens (inactive pairs skipped), with and
    without the side stream."""
    specs = [channel_spec(256, 0.002, 808e-9, 0.05, 4e3, 5, 256, 2e-15, pupils=[(0.05, 0, 0)]),
             channel_spec(256, 0.0015, 1550e-9, 0.04, 2e3, 2, 256, 5e-15, pupils=[(0.04, 0, 0)])]
    tables = engine.DeviceTables(specs, dev)
    chan = np.array([0, 1, 1, 0, 1])
    ns = np.array([specs[c].nscreens for c in chan])
    outs = []
    for whole in (True, False):
        monkeypatch.setattr(engine, "SYNTH_WHOLE_BATCH", whole)
        ws = engine.Workspace(256, 256, 5, "f16x3", dev)
        out = torch.empty(5, 256, 256, dtype=torch.complex64, device=dev)
        m = engine.run_batch(ws, tables, chan, ns, seed=13, rid=np.arange(5) + 40, field_out=out, overlap=overlap)
        torch.cuda.synchronize()
        assert engine.StepRunner(ws, tables, chan, ns).whole_batch_synthesis()
        outs.append((out.cpu().numpy(), m.cpu().numpy()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1][:, :8], outs[1][1][:, :8])


def test_ragged_batch_with_vacuum_element(dev):
    """A batch whose elements have 0 (vacuum), 2 and 5 screens: the vacuum
    element equals the FP64 vacuum propagation of the source over the whole
    length, the others equal their single-element runs, and every element's
    moments are finite."""
    s5 = channel_spec(128, 0.004, 808e-9, 0.05, 3e3, 5, 64, 1e-15)
    s2 = channel_spec(128, 0.003, 1550e-9, 0.04, 2e3, 2, 64, 3e-15)
    vac = engine.ChannelSpec(n=128, delta=0.004, wvl=808e-9, w0=0.05, F0=np.inf, length=3e3,
                             positions=np.zeros(0), points=64, base=s5.base, psd=s5.psd)
    specs = [vac, s2, s5]
    tables = engine.DeviceTables(specs, dev)
    chan = np.array([0, 1, 2, 0])
    ns = np.array([specs[c].nscreens for c in chan])
    ws = engine.Workspace(128, 64, 4, "f16x3", dev)
    out = torch.empty(4, 128, 128, dtype=torch.complex64, device=dev)
    mom = engine.run_batch(ws, tables, chan, ns, seed=3, rid=np.arange(4), field_out=out, reduce=True)
    assert torch.isfinite(mom[:, :7]).all()
    ref = o.vacuum_step(o.gaussian_source(128, 0.004, 808e-9, 0.05, np.inf), 0.004, 808e-9, 3e3)
    for b in (0, 3):
        assert rel_l2(out[b].cpu().numpy(), ref) < 2e-6, b
    for b in (1, 2):
        t1 = engine.DeviceTables([specs[chan[b]]], dev)
        w1 = engine.Workspace(128, 64, 1, "f16x3", dev)
        o1 = torch.empty(1, 128, 128, dtype=torch.complex64, device=dev)
        # the batch holds a vacuum element, so it takes the 2-D first step
        engine.run_batch(w1, t1, np.zeros(1), ns[b:b + 1], seed=3, rid=np.array([b]), field_out=o1, reduce=False,
                         separable=False)
        assert torch.equal(out[b], o1[0]), b
