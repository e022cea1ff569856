"""Measuring tuner for tile shape and launch shape.

The reference picks its chunk size by interpolating a table of VRAM sizes
(``config/auto_tune.py:57-60,132-134``) and never measures.  Here the
candidates are timed on the job itself, on the GPU, and the fastest is kept:

* ``tune_tile_size`` -- the core tile size of the tiled scheduler
  (``scheduler.process_dem``).  The halo is not a free parameter: it is the
  reference's padding rule for the algorithm (``params.required_padding``),
  since a narrower halo would change results and a wider one only costs.
* ``fujishadergpu_amd.batch.tune`` -- workgroups per CU and item order of the
  fused persistent batch kernel.

Results are cached per process by a caller-chosen key.
"""
from __future__ import annotations

import time
from typing import Callable, Dict, Hashable, Optional, Sequence, Tuple

import torch

_CACHE: Dict[Hashable, Tuple[int, Dict[int, float]]] = {}


def _gpu_time(fn: Callable[[], None], reps: int) -> float:
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(1, reps)):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / max(1, reps)


def tune_tile_size(run: Callable[[int], None], candidates: Sequence[int] = (2048, 4096, 8192),
                   reps: int = 1, key: Optional[Hashable] = None
                   ) -> Tuple[int, Dict[int, float]]:
    """Time ``run(tile_size)`` for each candidate (after one warm-up call) and
    return (best tile size, {tile size: seconds})."""
    if key is not None and key in _CACHE:
        return _CACHE[key]
    times: Dict[int, float] = {}
    for ts in candidates:
        run(int(ts))
        times[int(ts)] = _gpu_time(lambda: run(int(ts)), reps)
    best = min(times, key=times.get)
    if key is not None:
        _CACHE[key] = (best, times)
    return best, times
