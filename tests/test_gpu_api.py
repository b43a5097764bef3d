"""pyAtmosphere-compatible API on the device: the reference notebook's vacuum
known answers and the turbulent path against the oracle with the same seed."""
import numpy as np
import pytest
import torch

import pyatmosphere_amd as pa
from oracle import fp64 as o
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture
def vacuum_channel(dev):
    # tests/itest_vacuum_propagation.ipynb cell 3
    return pa.Channel(grid=pa.RectGrid(resolution=2 ** 8, delta=0.002),
                      source=pa.GaussianSource(wvl=809e-9, w0=0.02, F0=np.inf),
                      path=pa.VacuumPath(length=4e3), pupil=pa.CirclePupil(radius=0.5))


def test_vacuum_full_eta(vacuum_channel):
    np.testing.assert_almost_equal(pa.measures.eta(vacuum_channel, pupil=False), 1, decimal=6)


def test_vacuum_norm_amplitude(vacuum_channel):
    u = vacuum_channel.run(pupil=False)
    actual = abs(u.to(torch.complex128).sum().item()) * vacuum_channel.grid.delta ** 2
    np.testing.assert_almost_equal(actual, np.sqrt(2 * np.pi) * vacuum_channel.source.w0, decimal=7)


def test_vacuum_norm_intensity(vacuum_channel):
    I = pa.measures.I(vacuum_channel, pupil=False)
    np.testing.assert_almost_equal(I.double().sum().item() * vacuum_channel.grid.delta ** 2, 1, decimal=5)


def test_vacuum_waist(vacuum_channel):
    ch = vacuum_channel
    zR = np.pi * ch.source.w0 ** 2 / ch.source.wvl
    desired = ch.source.w0 * np.sqrt(1 + (ch.path.length / zR) ** 2)
    mean_r2 = pa.measures.mean_x2(ch) + pa.measures.mean_y2(ch)
    np.testing.assert_almost_equal(np.sqrt(2 * mean_r2), desired, decimal=7)


def _quick(n=256, delta=0.002, L=5e3, S=3):
    return pa.Channel(
        grid=pa.RectGrid(resolution=n, delta=delta),
        source=pa.GaussianSource(wvl=808e-9, w0=0.05, F0=np.inf),
        path=pa.IdenticalPhaseScreensPath(
            phase_screen=pa.SSPhaseScreen(model=pa.MVKModel(Cn2=1e-15, l0=6e-3, L0=1e3),
                                          f_grid=pa.RandLogPolarGrid(points=128, f_min=1 / 15000, f_max=2 / 6e-3)),
            length=L, count=S),
        pupil=pa.CirclePupil(radius=0.06))


def _oracle_for(ch, seed):
    rs = np.random.RandomState(seed)
    ps = ch.path.phase_screens[0]
    spectra = [o.draw_spectrum_numpy(rs, ps.f_grid.base, ps._get_psd()) for _ in ch.path.phase_screens]
    u0 = o.gaussian_source(ch.grid.resolution[0], ch.grid.delta, ch.source.wvl, ch.source.w0, ch.source.F0)
    return o.run_channel(u0, ch.grid.delta, ch.source.wvl, ch.path.length, ch.path.positions, spectra)


def test_run_matches_oracle_same_seed(dev):
    ch = _quick()
    np.random.seed(0)
    u = ch.run(pupil=False)
    ref = _oracle_for(ch, 0)
    assert rel_l2(u.cpu().numpy(), ref) < 1e-4
    np.random.seed(0)
    up = ch.run()
    mask = o.pupil_mask(256, ch.grid.delta, 0.06)
    assert rel_l2(up.cpu().numpy(), ref * mask) < 1e-4
    assert np.isclose(pa.measures.eta(ch, output=up), o.pupil_eta(ref, ch.grid.delta, 0.06), rtol=1e-4)


def test_generator_matches_run(dev):
    ch = _quick()
    np.random.seed(3)
    ref = ch.run(pupil=False)
    np.random.seed(3)
    steps = list(ch.generator(pupil=False))
    assert len(steps) == 3
    assert all(u.shape == (256, 256) and phi.shape == (256, 256) for u, phi in steps)
    assert torch.equal(ch.output, ref)
    # the field after each screen (pathes.py:61-75) against the oracle stepped
    # screen by screen with the same draws
    rs = np.random.RandomState(3)
    ps = ch.path.phase_screens[0]
    spectra = [o.draw_spectrum_numpy(rs, ps.f_grid.base, ps._get_psd()) for _ in range(3)]
    u = o.gaussian_source(256, 0.002, 808e-9, 0.05, np.inf).astype(np.complex128)
    for k, d in enumerate(o.step_lengths(5e3, ch.path.positions)[:3]):
        u = np.exp(-1j * o.screen_phase(256, 0.002, spectra[k])) * o.vacuum_step(u, 0.002, 808e-9, d)
        assert rel_l2(steps[k][0].cpu().numpy(), u) < 1e-4


def test_screen_generate_matches_oracle(dev):
    ch = _quick()
    ps = ch.path.phase_screens[0]
    ch.path.init_phase_screens()
    np.random.seed(7)
    phi = ps.generate()
    rs = np.random.RandomState(7)
    ref = o.screen_phase(256, ch.grid.delta, o.draw_spectrum_numpy(rs, ps.f_grid.base, ps._get_psd()))
    assert np.abs(phi.cpu().numpy() - ref).max() < 1e-4 + 2e-7 * np.abs(ref).max()


def test_measures_match_oracle(dev):
    ch = _quick()
    np.random.seed(1)
    u = ch.run(pupil=False)
    m = o.moments(u.cpu().numpy().astype(np.complex128), ch.grid.delta)
    for name in ["mean_x", "mean_y", "mean_x2", "mean_xy", "mean_y2"]:
        assert np.isclose(getattr(pa.measures, name)(ch, output=u), m[name], rtol=1e-9, atol=1e-15)


def test_use_gpu_false_raises(dev):
    pa.gpu.config["use_gpu"] = False
    try:
        with pytest.raises(RuntimeError, match="GPU only"):
            _quick().run()
    finally:
        pa.gpu.config["use_gpu"] = True


def test_channel_plan_matches_engine_and_oracle(dev):
    """pa_channel_create / pa_channel_run (the channel-level C entry points a
    non-Python caller binds) on the 512^2 channel with host draws: equal to the
    package's engine.run_batch on the same draws within 1e-6 (the plan builds
    the same float64 tables), within 1e-4 of the FP64 oracle; with device draws
    equal to run_batch with the same seed and realization ids."""
    import ctypes
    from pyatmosphere_amd import _lib, engine
    from tests.helpers import channel_spec, host_spectra, oracle_field
    spec = channel_spec(n=512, delta=0.0015, wvl=808e-9, w0=0.09, L=10e3, S=5, P=1024, Cn2=1e-15,
                        pupils=[(0.12, 0.0, 0.0), (0.05, 0.01, -0.02)])
    B = 4
    lib = _lib.load()
    plan = ctypes.c_void_p()
    pos = np.ascontiguousarray(spec.positions, np.float64)
    pup = np.ascontiguousarray(spec.pupils, np.float64)
    base = np.ascontiguousarray(spec.base, np.float32)
    psd = np.ascontiguousarray(spec.psd, np.float32)
    stream = _lib.stream_handle()
    _lib.check(lib.pa_channel_create(ctypes.byref(plan), spec.n, spec.delta, spec.wvl, spec.w0, np.inf, spec.length,
                                     spec.nscreens, pos.ctypes.data, None, spec.points, base.ctypes.data,
                                     psd.ctypes.data, len(spec.pupils), pup.ctypes.data, 0.0, 0.0, B, 0, stream),
               "pa_channel_create")
    try:
        screens, per_real = host_spectra(spec, B, seed=3)
        rho = torch.from_numpy(np.stack([s[0] for s in screens])).to(dev)
        theta = torch.from_numpy(np.stack([s[1] for s in screens])).to(dev)
        value = torch.from_numpy(np.stack([s[2] for s in screens])).to(dev)
        out = torch.empty(B, 512, 512, dtype=torch.complex64, device=dev)
        mom = torch.empty(B, _lib.PA_NUM_MOMENTS, dtype=torch.float64, device=dev)
        _lib.check(lib.pa_channel_run(plan, B, rho.data_ptr(), theta.data_ptr(), value.data_ptr(), None, 0, None,
                                      out.data_ptr(), mom.data_ptr(), stream), "pa_channel_run")
        tables = engine.DeviceTables([spec], dev)
        ws = engine.Workspace(512, 1024, B, "f16x3", dev)
        ref_out = torch.empty_like(out)
        ref_mom = engine.run_batch(ws, tables, np.zeros(B), np.full(B, 5), host_spectra=screens, field_out=ref_out,
                                   overlap=False)
        torch.cuda.synchronize()
        # the tables are float64 expressions rounded to complex64 on either side
        # (C++ std::exp here, numpy there): equal but for rare last-place
        # differences of single factors
        d = rel_l2(out.cpu().numpy(), ref_out.cpu().numpy())
        print(f"channel plan vs engine: field rel-L2 {d:.2e}, bit-equal {bool(torch.equal(out, ref_out))}", flush=True)
        assert d < 1e-6
        assert np.allclose(mom[:, :9].cpu().numpy(), ref_mom[:, :9].cpu().numpy(), rtol=1e-6, atol=1e-12)
        for b in range(B):
            assert rel_l2(out[b].cpu().numpy(), oracle_field(spec, per_real[b])) < 1e-4
        # device draws
        rid = torch.arange(100, 100 + B, dtype=torch.int64, device=dev)
        _lib.check(lib.pa_channel_run(plan, B, None, None, None, rid.data_ptr(), 77, None, None, mom.data_ptr(),
                                      stream), "pa_channel_run")
        ref_mom = engine.run_batch(ws, tables, np.zeros(B), np.full(B, 5), seed=77, rid=np.arange(100, 100 + B),
                                   overlap=False)
        torch.cuda.synchronize()
        assert np.allclose(mom[:, :9].cpu().numpy(), ref_mom[:, :9].cpu().numpy(), rtol=1e-6, atol=1e-12)
    finally:
        lib.pa_channel_destroy(plan)
