"""Recorded steps: the device work of a step captured once into a hipGraph
and replayed.

A step of the hot path for fixed shapes and buffers (the G_00 transform, the
shell x/y synthesis launches and the fused contractions of every angular
configuration) is a fixed sequence of kernel launches.  ``StepGraph``
records it with ``torch.cuda.CUDAGraph`` (a hipGraph on ROCm) on a side
stream of default priority, with the library context bound to that stream:

* the step runs once eagerly first, so FFT plans, the context's work areas
  and the autotuned launch shapes exist and the recording repeats the eager
  launches (the autotuner and the HBM block sizing replay their choices);
* calls that would create a plan, grow a work area or read back to the host
  fail while recording (the library checks the stream's capture status);
* the graph's private memory pool keeps every buffer the step allocates at
  the same address for all replays, and the context is held
  (``cnpc_ctx_hold``), so its work areas are not freed under the graph;
* ``replay()`` launches the whole step on torch's current stream, the one
  the context works on, so results are read as after the eager step.

Nothing recorded clears memory with ``hipMemsetAsync``: every kernel of the
step overwrites its outputs.
"""
import torch

from . import autotune, estimators
from ._lib import check


class StepGraph:
    """One recorded step.  ``record(enqueue, finish)`` returns what
    ``enqueue()`` returned while recording (typically the estimators'
    ``finish`` functions, whose device results every replay rewrites);
    ``finish(result)`` completes the eager run made before recording (by
    default a callable result is called)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.graph = None
        self._held = False

    def record(self, enqueue, finish=None, warmup=True):
        if self.graph is not None:
            raise RuntimeError("StepGraph already holds a recording")
        dev = self.ctx.device
        if warmup:
            done = enqueue()
            if finish is not None:
                finish(done)
            elif callable(done):
                done()
            del done
        torch.cuda.synchronize(dev)
        prev = torch.cuda.current_stream(dev)
        side = torch.cuda.Stream(dev)
        side.wait_stream(prev)
        g = torch.cuda.CUDAGraph()
        check(self.ctx.lib.cnpc_ctx_hold(self.ctx.h, 1), "cnpc_ctx_hold")
        self._held = True
        estimators._RECORDING[0] = True
        try:
            with autotune.frozen(), torch.cuda.graph(g, stream=side, capture_error_mode="thread_local"):
                self.ctx.bind_stream(side)
                try:
                    out = enqueue()
                finally:
                    self.ctx.bind_stream(prev)
        except BaseException:
            self.close()
            raise
        finally:
            estimators._RECORDING[0] = False
            self.ctx.bind_stream(prev)
        self.graph = g
        return out

    def replay(self):
        if self.graph is None:
            raise RuntimeError("StepGraph: nothing recorded")
        self.graph.replay()

    def close(self):
        """Drop the recording and its memory pool and release the hold."""
        if self.graph is not None:
            torch.cuda.synchronize(self.ctx.device)
            self.graph.reset()
            self.graph = None
        if self._held:
            self._held = False
            check(self.ctx.lib.cnpc_ctx_hold(self.ctx.h, -1), "cnpc_ctx_hold")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
