"""Launch-shape autotuner.

A few launch parameters of the hot path have no single best value across
grid sizes and bin counts: the number of shells per synthesis launch (more
shells amortise the shared cube and the launches, fewer keep the work area
small) and the number of split-K workgroups of the separate-path GEMM. The
first time a shape is seen, each candidate is timed on the real operation
(HIP events on the context stream, after one untimed call) and the fastest
is kept for the rest of the process; with ``CNPC_TUNE_CACHE`` set to a file
path the choices also persist across processes (JSON).

Every candidate computes the same result, so tuning never changes numbers.
``CNPC_TUNE=0`` disables tuning (first candidate), and so does ``frozen()``
(used while a graph is recorded: timing would synchronise the stream).
"""
import contextlib
import json
import os

import torch

_CHOICES = {}
_LOADED = False
_FROZEN = [False]


@contextlib.contextmanager
def frozen():
    """Within this block ``choose`` times nothing: cached choices or the
    first candidate."""
    prev = _FROZEN[0]
    _FROZEN[0] = True
    try:
        yield
    finally:
        _FROZEN[0] = prev


def _cache_path():
    return os.environ.get("CNPC_TUNE_CACHE")


def _load():
    global _LOADED
    if _LOADED:
        return
    _LOADED = True
    path = _cache_path()
    if path and os.path.exists(path):
        try:
            with open(path) as f:
                _CHOICES.update(json.load(f))
        except (OSError, ValueError):
            pass


def _save():
    path = _cache_path()
    if not path:
        return
    try:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, "w") as f:
            json.dump(_CHOICES, f, indent=1, sort_keys=True)
    except OSError:
        pass


def choices():
    """The tuned values so far, {key: value}."""
    _load()
    return dict(_CHOICES)


def lookup(name, shape):
    """The cached choice for this shape, or None."""
    _load()
    return _CHOICES.get(f"{name}:{shape}")


def choose(name, shape, candidates, run, units=None):
    """Fastest of ``candidates`` for ``run(candidate)`` at this ``shape``.
    ``units(candidate)`` normalises the time (e.g. per shell); the result is
    cached under name + shape."""
    _load()
    key = f"{name}:{shape}"
    if key in _CHOICES:
        return _CHOICES[key]
    if len(candidates) == 1 or _FROZEN[0] or os.environ.get("CNPC_TUNE", "1") == "0":
        return candidates[0]
    best, best_t = candidates[0], float("inf")
    for c in candidates:
        run(c)  # untimed: plans, work areas
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        run(c)
        e1.record()
        e1.synchronize()
        t = e0.elapsed_time(e1) / (units(c) if units else 1.0)
        if t < best_t:
            best, best_t = c, t
    _CHOICES[key] = best
    _save()
    return best
