"""Launch shapes on the device: every shape gives the same screen, and the
autotuner's choices are valid and picked up by new workspaces."""
import numpy as np
import pytest
import torch

from pyatmosphere_amd import autotune, engine
from tests.helpers import channel_spec

pytestmark = pytest.mark.gpu


@pytest.fixture
def cache(tmp_path, monkeypatch):
    monkeypatch.setenv("PA_TUNE_CACHE", str(tmp_path / "tune.json"))


def _screen_factor(ws, tables, B, tile, rows):
    st = torch.cuda.current_stream().cuda_stream
    engine.prep_screen(ws, tables, 0, B, None, st)
    engine.plane_wave_tables(ws, tables, 0, B, None, st, rows=rows)
    engine.screen_gemm(ws, 0, B, None, st, ang=ws.ang, tile=tile)
    return ws.ang[:B].cpu().numpy().copy()


@pytest.mark.parametrize("n", [512, 2048])
def test_launch_shapes_give_the_same_screen(dev, n):
    spec = channel_spec(n, 0.0015, 808e-9, 0.12, 50e3, 5, 1024, 5e-16)
    tables = engine.DeviceTables([spec], dev)
    B = 2
    ws = engine.Workspace(n, 1024, B, "f16x3", dev)
    ws.chan[:B].zero_()
    ws.rid[:B].copy_(torch.arange(B, dtype=torch.int64))
    engine.draw_device(ws, tables, 99, 0, B, torch.cuda.current_stream().cuda_stream)
    ref = _screen_factor(ws, tables, B, 1, 16)
    for tile in (1, 2):
        for rows in autotune.TABLE_ROWS:
            got = _screen_factor(ws, tables, B, tile, rows)
            err = np.abs(got - ref).max()
            print(f"n={n} tile={tile} rows={rows}: max |diff| {err:.2e}", flush=True)
            # the table recurrence restarts at a different row for each rows
            # setting (float64, then rounded to the split operands)
            assert err < 2e-6


def test_tune_screen_choice_is_used_by_new_workspaces(dev, cache):
    e = autotune.tune_screen(512, 256, 2, "f16x3", dev, reps=2)
    assert e["gemm_tile"] in autotune.GEMM_TILES and e["table_rows"] in autotune.TABLE_ROWS
    ws = engine.Workspace(512, 256, 2, "f16x3", dev)
    assert (ws.gemm_tile, ws.table_rows) == (e["gemm_tile"], e["table_rows"])


def test_tune_batch_picks_a_measured_candidate(dev, cache):
    e = autotune.tune_batch(256, 64, 2, "f16x3", dev, candidates=(1, 4), steps=2)
    assert e["batch"] in (1, 4) and len(e["rates"]) == 4
    assert engine.default_batch_for(256, 64, 2, "f16x3", dev) == e["batch"]
    ws = engine.Workspace(256, 64, e["batch"], "f16x3", dev)
    assert engine.use_overlap(ws, e["batch"], 2) == e["overlap"]


@pytest.mark.parametrize("n", [1024, 2048, 4096])
def test_column_groups_give_the_same_field(dev, n):
    """Every columns-per-block launch shape of the column pass computes the
    same column transforms: bit-equal fields after a full Fresnel step."""
    from pyatmosphere_amd import _lib
    spec = channel_spec(n, 0.0015, 808e-9, 0.12, 50e3, 5, 64, 5e-16)
    tables = engine.DeviceTables([spec], dev)
    B = 2
    rng = np.random.default_rng(n)
    u0 = torch.from_numpy((rng.normal(size=(B, n, n)) + 1j * rng.normal(size=(B, n, n))).astype(np.complex64)).to(dev)
    tw = engine.twiddles(n, dev)
    st = torch.cuda.current_stream().cuda_stream
    outs = []
    groups = autotune.col_groups(n)
    assert len(groups) >= 2
    for c in groups:
        work = torch.empty(B, n, n, dtype=torch.complex64, device=dev)
        out = torch.empty_like(work)
        _lib.call("pa_row_first", n, work.data_ptr(), u0.data_ptr(), None, None, tw.data_ptr(), B, st)
        _lib.call("pa_col_step", n, work.data_ptr(), tables.htab.data_ptr(), None, None, 1, tables.max_steps,
                  tw.data_ptr(), c, B, st)
        _lib.call("pa_row_last", n, work.data_ptr(), out.data_ptr(), None, None, None, 0, None, None, tw.data_ptr(),
                  B, st)
        outs.append(out.cpu())
    for c, o in zip(groups[1:], outs[1:]):
        assert torch.equal(o, outs[0]), c
