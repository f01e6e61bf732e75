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
        dev = torch.cuda.get_device_properties(torch.cuda.current_device())
        return json.dumps([getattr(dev, "gcnArchName", dev.name), *key])

    def best(self, key, candidates, time_fn):
        k = self._key(key)
        if k in self.table and os.environ.get("VCT_RETUNE") != "1":
            return self.table[k]
        if torch.cuda.is_current_stream_capturing():
            raise RuntimeError("autotuning inside HIP-graph capture")
        results = []
        for cfg in candidates():
            ms = float(time_fn(cfg))
            results.append((ms, cfg))
        results.sort(key=lambda r: r[0])
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
