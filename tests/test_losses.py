"""Path losses (pathes.py:19-24, 67-75): the per-segment amplitude factors the
device path folds into its Fresnel steps, against the reference semantics
restated by the oracle, and the oracle's factors against what a loss in dB
means (power falls by 10^(-dB/10))."""
import numpy as np
import pytest

import pyatmosphere_amd as pa
from oracle import fp64 as o


def _path(position, losses_db, L=5e3, S=4):
    path = pa.IdenticalPhaseScreensPath(
        phase_screen=pa.SSPhaseScreen(model=pa.MVKModel(Cn2=1e-15, l0=6e-3, L0=1e3),
                                      f_grid=pa.RandLogPolarGrid(points=16, f_min=1 / 15000, f_max=2 / 6e-3)),
        length=L, count=S, position_in_slab=position, losses_db=losses_db)
    return path


def test_append_losses_is_a_power_loss_in_db():
    u = np.full((4, 4), 2.0 + 1.0j)
    for db in (0.5, 3.0, 10.0):
        out = o.append_losses(u, db, 7.0)
        assert np.allclose(np.abs(out) ** 2, np.abs(u) ** 2 * 10 ** (-db / 10), rtol=1e-14)
    # `losses_db or path_losses_db`: a zero-dB segment takes the whole path's loss
    assert np.allclose(o.append_losses(u, 0.0, 6.0), u * 10 ** (-6 / 20))
    assert np.array_equal(o.append_losses(u, 0.0, 0.0), u)


@pytest.mark.parametrize("position", ["before", "middle", "after"])
@pytest.mark.parametrize("losses_db", [0.0, 3.0])
def test_segment_gains_follow_the_reference(position, losses_db):
    """Gain of Fresnel step i (the device applies segment i's loss with the step
    that precedes screen i; the factors commute) equals the oracle's amplitude
    ratio for that segment; the last step carries the path total only in
    output() (pathes.py:23-24), never in lossless_output() or the generator."""
    path = _path(position, losses_db)
    ds = o.step_lengths(path.length, path.positions)
    one = np.ones(1, np.complex128)
    want = [o.append_losses(one, losses_db * d / path.length, losses_db)[0].real for d in ds[:-1]]
    for with_total, last in ((False, 1.0), (True, o.append_losses(one, None, losses_db)[0].real)):
        g = path._segment_gains(with_total)
        assert len(g) == len(ds)
        assert np.allclose(g[:-1], want, rtol=1e-15)
        assert np.isclose(g[-1], last, rtol=1e-15)
    if position == "before" and losses_db:
        # the zero-length first step takes the whole path's loss (the quirk of
        # `losses_db or self.losses_db`, pathes.py:20)
        assert ds[0] == 0.0 and np.isclose(path._segment_gains(False)[0], 10 ** (-losses_db / 20))


@pytest.mark.gpu
@pytest.mark.parametrize("position", ["before", "middle"])
def test_outputs_with_losses_match_oracle(dev, position):
    """output, lossless_output and the generator with losses_db = 3 dB against
    the oracle with the same draws (np.random seed), rel-L2 < 1e-4."""
    from tests.helpers import rel_l2
    n, delta, L, S, db = 256, 0.002, 5e3, 3, 3.0
    ch = pa.Channel(grid=pa.RectGrid(resolution=n, delta=delta),
                    source=pa.GaussianSource(wvl=808e-9, w0=0.05, F0=np.inf),
                    path=_path(position, db, L=L, S=S), pupil=pa.CirclePupil(radius=0.06))
    ps = ch.path.phase_screens[0]
    ch.path.init_phase_screens()
    u0 = o.gaussian_source(n, delta, 808e-9, 0.05, np.inf)

    def oracle(seed, total):
        rs = np.random.RandomState(seed)
        spectra = [o.draw_spectrum_numpy(rs, ps.f_grid.base, ps._get_psd()) for _ in range(S)]
        return o.run_channel(u0, delta, 808e-9, L, ch.path.positions, spectra, losses_db=db,
                             total_losses=total, return_steps=True)

    np.random.seed(11)
    out = ch.run(pupil=False)
    ref, _ = oracle(11, True)
    assert rel_l2(out.cpu().numpy(), ref) < 1e-4
    np.random.seed(12)
    src = ch.source.output()
    lossless = ch.path.lossless_output(src)
    ref, _ = oracle(12, False)
    assert rel_l2(lossless.cpu().numpy(), ref) < 1e-4
    np.random.seed(13)
    steps = list(ch.generator(pupil=False))
    ref, ref_steps = oracle(13, False)
    for (u, _), r in zip(steps, ref_steps):
        assert rel_l2(u.cpu().numpy(), r) < 1e-4
    assert rel_l2(ch.output.cpu().numpy(), ref) < 1e-4
