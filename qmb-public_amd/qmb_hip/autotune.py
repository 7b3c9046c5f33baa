"""
Launch-shape autotuner.

Times candidate launch parameters of a kernel call with HIP events on the current stream and keeps
the fastest per shape signature.  The cache lives in memory and, when QMB_HIP_TUNE_CACHE names a
file, in that JSON file as well, so a later process starts from the tuned shapes.  Tuning runs the
real kernels on the caller's data into scratch outputs; it synchronises the device.
"""

from __future__ import annotations

import json
import os
import pathlib
import statistics
import typing

import torch


class Autotuner:
    """Fastest candidate per key: {key: {"params": ..., "ms": ..., "tried": [[params, ms], ...]}}."""

    def __init__(self, path: str | None = None) -> None:
        self.path = pathlib.Path(path) if path else None
        self.cache: dict[str, dict[str, typing.Any]] = {}
        if self.path is not None and self.path.exists():
            try:
                self.cache = json.loads(self.path.read_text())
            except (OSError, ValueError):
                self.cache = {}

    def best(self, key: str, candidates: typing.Sequence[dict[str, typing.Any]], run: typing.Callable[[dict[str, typing.Any]], None],
             reps: int = 3, warmup: int = 1, retune: bool = False) -> dict[str, typing.Any]:
        """Parameters of the fastest `run(params)` among `candidates` (cached under `key`)."""
        if not retune and key in self.cache:
            return dict(self.cache[key]["params"])
        if not candidates:
            raise ValueError("no candidates to tune")
        stream = torch.cuda.current_stream()
        tried = []
        for params in candidates:
            for _ in range(warmup):
                run(params)
            times = []
            for _ in range(reps):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                run(params)
                b.record(stream)
                b.synchronize()
                times.append(a.elapsed_time(b))
            tried.append((dict(params), statistics.median(times)))
        params, ms = min(tried, key=lambda x: x[1])
        self.cache[key] = {"params": params, "ms": ms, "tried": [[p, t] for p, t in tried]}
        self.save()
        return dict(params)

    def save(self) -> None:
        if self.path is not None:
            self.path.write_text(json.dumps(self.cache, indent=1, sort_keys=True))


TUNER = Autotuner(os.environ.get("QMB_HIP_TUNE_CACHE"))
