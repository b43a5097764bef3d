"""Monte Carlo driver (simulations/simulation.py:7-152 of pyAtmosphere).

Measures are grouped channel -> time -> measure type -> operations as in the
reference.  When every measure of a run is one of the standard end-of-path
observables (measures.mean_x/mean_y/mean_x2/mean_xy/mean_y2/eta,
BeamResult.mean_x2_r, the PDTResult and TrackedPDTResult pupil chains), the
realizations are run in batches through the fused device engine: one launch
sequence propagates many realizations and one reduction pass produces every
moment and pupil transmittance, so no field is materialised per measure.
Otherwise each realization runs through Channel.generator and the measures'
own operations, exactly as the reference loop does.

Spectra come from the host legacy numpy stream in the reference draw order
(rng="numpy", the default: np.random.seed makes runs reproducible) or from
the device Philox stream (rng="device", `seed`), which avoids host work.
"""
from functools import partial
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from .. import engine, measures as M
from ..gpu import config, device
from ..pathes import PhaseScreensPath
from ..phase_screens import SSPhaseScreen
from ..sources import GaussianSource
from .beam import BeamResult
from .measure import Measure
from .pdt import PDTResult, TrackedPDTResult
from .result import Result

_MOMENT_OPS = {M.eta: 0, M.mean_x: 1, M.mean_y: 2, M.mean_x2: 3, M.mean_xy: 4, M.mean_y2: 5}


def _classify(measure):
    """('moment', index) | ('pupil', radius, sx, sy) | ('tracked', radius) | None."""
    if measure.time is not None or measure.measure_type != "atmosphere":
        return None
    ops = measure.operations
    if len(ops) == 1:
        op = ops[0]
        if op in _MOMENT_OPS:
            return ("moment", _MOMENT_OPS[op])
        if getattr(op, "__func__", None) is BeamResult.mean_x2_r:
            return ("moment", 6)
        return None
    if len(ops) == 2 and ops[1] is M.eta and isinstance(ops[0], partial) \
            and getattr(ops[0].func, "__func__", None) is PDTResult.append_pupil:
        res = ops[0].func.__self__
        pupil = ops[0].args[0]
        if isinstance(res, TrackedPDTResult):
            return None
        return ("pupil", float(pupil.radius), float(res.pupil_shift[0]), float(res.pupil_shift[1]))
    if len(ops) == 3 and ops[2] is M.eta and getattr(ops[0], "__func__", None) is TrackedPDTResult.set_pupil_position \
            and isinstance(ops[1], partial):
        return ("tracked", float(ops[1].args[0].radius))
    return None


default_batch = engine.default_batch_for


class Simulation:
    def __init__(self, results_list: Sequence[Result] = None, measures_list: Sequence[Measure] = None,
                 batch: Optional[int] = None, rng: str = "numpy", seed: int = 0):
        self.measures: Dict = {}
        if measures_list:
            for m in measures_list:
                self.add_measures(m)
        self.results_list = results_list
        if results_list:
            for result in results_list:
                for m in result.measures:
                    self.add_measures(m)
        if rng not in ("numpy", "device"):
            raise ValueError("rng must be 'numpy' or 'device'")
        self.batch, self.rng, self.seed = batch, rng, seed
        self._rid: Dict = {}

    # ------------------------------------------------------------ reference bookkeeping
    def add_measures(self, measures):
        """channel - time - measure_type - operations"""
        d = self.measures.setdefault(measures.channel, {})
        d = d.setdefault(measures.time, {})
        d = d.setdefault(measures.measure_type, {})
        d.setdefault(measures.operations, []).append(measures)

    def init_measures_iteration_data(self):
        for m in self.flattened_measures():
            empty = None
            npos = len(m.channel.path.positions) + 1 if hasattr(m.channel.path, "positions") else 1
            if m.time:
                if m.measure_type == "propagation":
                    empty = [[None] * npos for _ in m.time]
                else:
                    empty = [None for _ in m.time]
            elif m.measure_type == "propagation":
                empty = [None] * npos
            m.iteration_data = empty

    def process_operations(self, output, operations_measures, time_id, propagation_id=None):
        for operations, measures_list in operations_measures.items():
            if self.is_measures_done(measures_list):
                continue
            first = measures_list[0]
            out = output.clone() if torch.is_tensor(output) else output
            for operation in operations:
                out = operation(first.channel, output=out)
            for m in measures_list:
                if m.is_done:
                    continue
                if m.time is not None:
                    if propagation_id is not None:
                        m.iteration_data[time_id][propagation_id] = out
                    else:
                        m.iteration_data[time_id] = out
                elif propagation_id is not None:
                    m.iteration_data[propagation_id] = out
                else:
                    m.iteration_data = out

    def iter(self):
        """One realization of every channel through Channel.generator (generic path)."""
        self.init_measures_iteration_data()
        for channel, channel_measures in self.measures.items():
            for ps in channel.path.phase_screens:
                ps.cache_clear()
            for time, time_measures in channel_measures.items():
                time = time or [None]
                for time_id, time_value in enumerate(time):
                    gen = channel.generator(pupil=False, shift=(0, time_value or 0), store_output=True, wind=True)
                    for propagation_id, (result, phase_screen) in enumerate(gen):
                        self.process_operations(result, time_measures.get("propagation", {}), time_id, propagation_id)
                        if propagation_id == 0:
                            self.process_operations(phase_screen, time_measures.get("phase_screen", {}), time_id)
                    self.process_operations(channel.output, time_measures.get("atmosphere", {}), time_id)
                    self.process_operations(channel.output, time_measures.get("propagation", {}), time_id, -1)
                    if channel.pupil:
                        self.process_operations(channel.pupil.output(channel.output),
                                                time_measures.get("pupil", {}), time_id)
        for m in self.flattened_measures():
            if not m.is_done:
                m.data.append(m.iteration_data)

    def flattened_measures(self, measures=None):
        measures = measures if measures is not None else self.measures
        if isinstance(measures, dict):
            for values in measures.values():
                yield from self.flattened_measures(values)
        else:
            yield from measures

    def is_measures_done(self, measures=None):
        return all(m.is_done for m in self.flattened_measures(measures))

    def run(self, *args, plot_step: int = None, save_step: int = None, **kwargs):
        try:
            plan = self._fused_plan()
            if plan is not None:
                self._run_fused(plan, plot_step, save_step)
            else:
                iteration = 0
                while not self.is_measures_done():
                    self.iter()
                    iteration += 1
                    self.process_output(iteration, plot_step=plot_step, save_step=save_step)
        except KeyboardInterrupt:
            pass
        finally:
            self.process_output(0, plot_step=plot_step, save_step=save_step)

    def process_output(self, iteration, plot_step, save_step):
        if plot_step and iteration % plot_step == 0:
            for result in self.results_list or []:
                result.plot_output()
            try:
                from IPython import display
                display.clear_output(wait=True)
            except ModuleNotFoundError:
                pass
        if save_step and iteration % save_step == 0:
            for result in self.results_list or []:
                result.save_output()

    # ------------------------------------------------------------ fused batched path
    def _fused_plan(self):
        plans = []
        for channel, by_time in self.measures.items():
            if set(by_time) != {None}:
                return None
            path = channel.path
            if not isinstance(path, PhaseScreensPath) or not isinstance(channel.source, GaussianSource):
                return None
            if not all(isinstance(ps, SSPhaseScreen) for ps in path.phase_screens):
                return None
            items = []
            pupils: List = []
            tracked: List = []
            for m in self.flattened_measures(by_time):
                c = _classify(m)
                if c is None:
                    return None
                if c[0] == "pupil":
                    key = c[1:]
                    if key not in pupils:
                        pupils.append(key)
                    c = ("pupil", pupils.index(key))
                elif c[0] == "tracked":
                    if c[1] not in tracked:
                        tracked.append(c[1])
                    c = ("tracked", tracked.index(c[1]))
                items.append((m, c))
            if len(pupils) > engine._lib.PA_MAX_PUPILS or len(tracked) > engine._lib.PA_MAX_PUPILS:
                return None
            plans.append(dict(channel=channel, items=items, pupils=pupils, tracked=tracked))
        return plans

    def _remaining(self):
        need = 0
        for m in self.flattened_measures():
            if m.max_size is None:
                return None
            need = max(need, m.max_size - len(m))
        return need

    def _run_fused(self, plans, plot_step, save_step):
        dev = device()
        mode = config.get("gemm_mode", "f16x3")
        prepared = []
        for p in plans:
            ch = p["channel"]
            path = ch.path
            path.init_phase_screens()
            spec = path._spec((0.0, 0.0), with_total=False)  # Channel.generator output
            spec.pupils = [(r, sx, sy) for (r, sx, sy) in p["pupils"]]
            tables = engine.DeviceTables([spec], dev)
            prepared.append((p, spec, tables))
        iteration = 0
        while not self.is_measures_done():
            remaining = self._remaining()
            B = self.batch or max(default_batch(s.n, s.points, s.nscreens, mode, dev) for _, s, _ in prepared)
            if remaining is not None:
                B = max(1, min(B, remaining))
            host = {}
            if self.rng == "numpy":
                # reference draw order: realization, then channel, then screen
                draws = {id(p): [] for p, _, _ in prepared}
                for _ in range(B):
                    for p, spec, _ in prepared:
                        per = []
                        for ps in p["channel"].path.phase_screens:
                            ps.cache_clear()
                            sp = ps._get_spectrum(use_cached_spectrum=True)
                            per.append(sp)
                        draws[id(p)].append(per)
                for p, spec, _ in prepared:
                    d = draws[id(p)]
                    host[id(p)] = [(np.stack([d[b][i].rho for b in range(B)]),
                                    np.stack([d[b][i].theta for b in range(B)]),
                                    np.stack([d[b][i].value for b in range(B)])) for i in range(spec.nscreens)]
            results = {}
            for p, spec, tables in prepared:
                ws = engine.get_workspace(spec.n, spec.points, B, mode, dev)
                rid = None
                if self.rng == "device":
                    start = self._rid.get(id(p["channel"]), 0)
                    rid = np.arange(start, start + B)
                    self._rid[id(p["channel"])] = start + B
                out = engine.run_batch(ws, tables, np.zeros(B), np.full(B, spec.nscreens), seed=self.seed, rid=rid,
                                       host_spectra=host.get(id(p)), tracked=p["tracked"] or None)
                # centroid-tracked pupils: a second reduction pass over the last
                # work buffer (no field is stored)
                mom, tr = (out[0].cpu().numpy(), out[1].cpu().numpy()) if p["tracked"] else (out.cpu().numpy(), None)
                results[id(p)] = (mom, tr)
            for b in range(B):
                for p, spec, _ in prepared:
                    mom, tr = results[id(p)]
                    for m, c in p["items"]:
                        if c[0] == "moment":
                            m.iteration_data = float(mom[b, c[1]])
                        elif c[0] == "pupil":
                            m.iteration_data = float(mom[b, 7 + c[1]])
                        else:
                            m.iteration_data = float(tr[b, 7 + c[1]])
                    for m, c in p["items"]:
                        if c[0] == "tracked":
                            res = m.operations[0].__self__
                            res.pupil_shift = (float(mom[b, 1]), float(mom[b, 2]))
                for m in self.flattened_measures():
                    if not m.is_done:
                        m.data.append(m.iteration_data)
                iteration += 1
                self.process_output(iteration, plot_step=plot_step, save_step=save_step)
                if self.is_measures_done():
                    break
