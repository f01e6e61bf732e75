"""Launch-shape autotuner.

For a key (kernel family + problem shape + device), every candidate
configuration is timed with HIP events on the current stream, the fastest is
kept in memory and in a JSON cache (``$VCT_TUNE_CACHE`` or
``~/.cache/virtualct_amd/tune.json``), so later processes reuse it.  Tuning
never runs during HIP-graph capture: callers tune in their constructors.
"""
from __future__ import annotations

import json
import os
import threading

import torch

_lock = threading.Lock()
_tuner = None


class Tuner:
    def __init__(self, path: str | None = None):
        self.path = path or os.environ.get(
            "VCT_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "virtualct_amd", "tune.json"))
        self.table: dict[str, dict] = {}
        self.log: list[dict] = []
        try:
            with open(self.path) as f:
                self.table = json.load(f)
        except (OSError, ValueError):
            self.table = {}

    @staticmethod
    def _key(key) -> str:
        if torch.cuda.is_available():
            dev = torch.cuda.get_device_properties(torch.cuda.current_device())
            name = getattr(dev, "gcnArchName", dev.name)
        else:
            name = "cpu"
        return json.dumps([name, *key])

    def best(self, key, candidates, time_fn):
        k = self._key(key)
        if k in self.table and os.environ.get("VCT_RETUNE") != "1":
            return self.table[k]
        if torch.cuda.is_available() and torch.cuda.is_current_stream_capturing():
            raise RuntimeError("autotuning inside HIP-graph capture")
        cands = list(candidates())
        # one untimed pass (clocks ramp up on the first launches after idle),
        # then the best of three interleaved passes per candidate
        for cfg in cands:
            time_fn(cfg)
        best_ms = [float("inf")] * len(cands)
        for _ in range(3):
            for i, cfg in enumerate(cands):
                best_ms[i] = min(best_ms[i], float(time_fn(cfg)))
        results = sorted(zip(best_ms, cands), key=lambda r: r[0])
        best = dict(results[0][1])
        self.table[k] = best
        self.log.append({"key": key, "results": [(ms, cfg) for ms, cfg in results]})
        self._save()
        return best

    def _save(self):
        try:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            tmp = self.path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self.table, f, indent=1)
            os.replace(tmp, self.path)
        except OSError:
            pass


def tuner() -> Tuner:
    global _tuner
    with _lock:
        if _tuner is None:
            _tuner = Tuner()
        return _tuner


def events_ms(fn, reps: int = 3) -> float:
    """Average milliseconds of fn() over `reps` calls after one untimed call,
    from HIP events on the current stream."""
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


TRACE_BLOCKS = (64, 128, 256)
ADJOINT_VIEWS = (0, 25, 50, 100, 200)
BP_SINOS_PER_LANE = (4, 8, 12, 16)


def tune_plan(plan, M: int = 6, n_sino: int = 1, what=("trace", "adjoint", "backproject")) -> dict:
    """Time the launch shapes of one plan on synthetic data of its geometry and
    set the fastest as plan options: the tracer's lanes per workgroup, the
    adjoint's views per workgroup and (n_sino >= 2) the sinograms per lane of
    an n_sino-sinogram backprojection.  Winners are cached per shape key;
    results do not depend on the choice beyond float rounding in the adjoint.
    Returns the chosen values."""
    from . import _lib, ops
    dev = torch.device("cuda", plan.device)
    g = plan.geom
    shape = (g.recon_size, round(float(g.pixel_cm), 9), g.n_views, g.n_channels)
    gen = torch.Generator(device=dev).manual_seed(1)
    chosen = {}
    if "trace" in what:
        dens = torch.randint(0, 2000, (M, plan.N, plan.N), generator=gen, device=dev, dtype=torch.int16)
        pk, dt = ops.pack_densities(plan, dens)

        def t_trace(cfg):
            plan.set_option(_lib.OPT_TRACE_BLOCK, cfg["trace_block"])
            return events_ms(lambda: ops.trace_paths(plan, pk, dt, M, check=False))
        cfg = tuner().best(("trace block", *shape, M), lambda: [{"trace_block": b} for b in TRACE_BLOCKS], t_trace)
        plan.set_option(_lib.OPT_TRACE_BLOCK, cfg["trace_block"])
        chosen.update(cfg)
    if "adjoint" in what:
        w = torch.zeros((plan.V, plan.W, 8), device=dev)
        w[..., :M] = torch.randn((plan.V, plan.W, M), generator=gen, device=dev)
        grad = torch.zeros((plan.N, plan.N, 8), device=dev)

        def t_adj(cfg):
            plan.set_option(_lib.OPT_ADJOINT_VIEWS, cfg["adjoint_views"])
            return events_ms(lambda: ops.trace_adjoint(plan, w, M, grad=grad, check=False))
        cfg = tuner().best(("adjoint views", *shape, M), lambda: [{"adjoint_views": v} for v in ADJOINT_VIEWS], t_adj)
        plan.set_option(_lib.OPT_ADJOINT_VIEWS, cfg["adjoint_views"])
        chosen.update(cfg)
    if "backproject" in what and n_sino >= 2:
        fl = torch.randn((n_sino, plan.V, plan.W), generator=gen, device=dev)

        def t_bp(cfg):
            plan.set_option(_lib.OPT_BP_SINOS_PER_LANE, cfg["bp_sinos_per_lane"])
            return events_ms(lambda: ops.backproject(plan, fl))
        cfg = tuner().best(("bp sinograms per lane", *shape, n_sino),
                           lambda: [{"bp_sinos_per_lane": k} for k in BP_SINOS_PER_LANE], t_bp)
        plan.set_option(_lib.OPT_BP_SINOS_PER_LANE, cfg["bp_sinos_per_lane"])
        chosen.update(cfg)
    ops.check_errors(plan.device)
    return chosen
