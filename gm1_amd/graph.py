"""The large step captured into a HIP graph and replayed (host mirror of cuda_graphs.ScheduleGraphs).

The reference captures one step's launches with CUDA graphs
(``cuda_graphs.ScheduleGraphs``, cuda_graphs.py:754, keyed by
``cuda_cartesian_runtime.graph_key`` :987-1020) to remove host launch cost.
Here every operator of ``launch_cuda_gm1_large_step`` enqueues on one stream
with no host synchronisation and no allocation, so ``steps`` consecutive
large steps are captured once with ``torch.cuda.CUDAGraph`` and replayed;
the replay reads and writes the same device buffers as the eager step.

Ordering: a replay runs on the graph's own stream.  Before it, that stream
waits for the caller's current stream and the state's bound stream (so an
upload or perturbation queued there lands first); after it, both of those
streams wait for the replay (so whatever the caller queues next sees the
advanced state).  No host synchronisation is involved.
"""

from __future__ import annotations

from typing import Any

import torch

from .large_step import Gm1LargeStepState, launch_cuda_gm1_large_step


class Gm1LargeStepGraph:
    """``steps`` large steps of ``state`` as one replayable HIP graph."""

    def __init__(self, state: Gm1LargeStepState, steps: int = 1, stream: Any | None = None):
        if steps < 1:
            raise ValueError("steps must be >= 1")
        self.state = state
        self.steps = steps
        self.graph = torch.cuda.CUDAGraph()
        self.stream = stream if stream is not None else (state.stream or torch.cuda.Stream())
        saved_stream, saved_index = state.stream, state.step_index
        torch.cuda.synchronize()
        try:
            state.stream = None  # launch on the capturing (current) stream
            with torch.cuda.graph(self.graph, stream=self.stream):
                for _ in range(steps):
                    launch_cuda_gm1_large_step(state)
        finally:
            state.stream, state.step_index = saved_stream, saved_index

    def _neighbours(self) -> list:
        out = [torch.cuda.current_stream()]
        if self.state.stream is not None and self.state.stream != out[0]:
            out.append(self.state.stream)
        return [s for s in out if s != self.stream]

    def replay(self, ordered: bool = True) -> None:
        """Enqueue the captured steps (asynchronously).

        ``ordered`` (the default) makes the replay wait for the caller's current stream and the state's
        stream, and makes them wait for the replay.  A caller that orders a batch of replays itself (the
        ensemble forks once before its members and joins once after) passes ``ordered=False``.
        """
        around = self._neighbours() if ordered else []
        for s in around:
            self.stream.wait_stream(s)
        with torch.cuda.stream(self.stream):
            self.graph.replay()
        for s in around:
            s.wait_stream(self.stream)
        self.state.step_index += self.steps


__all__ = ["Gm1LargeStepGraph"]
