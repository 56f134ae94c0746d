"""The large step captured into a HIP graph and replayed (host mirror of cuda_graphs.ScheduleGraphs).

The reference captures one step's launches with CUDA graphs
(``cuda_graphs.ScheduleGraphs``, cuda_graphs.py:754, keyed by
``cuda_cartesian_runtime.graph_key`` :987-1020) to remove host launch cost.
Here every operator of ``launch_cuda_gm1_large_step`` enqueues on one stream
with no host synchronisation and no allocation, so ``steps`` consecutive
large steps are captured once with ``torch.cuda.CUDAGraph`` and replayed;
the replay reads and writes the same device buffers as the eager step.
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

    def replay(self) -> None:
        """Enqueue the captured steps (asynchronously) on the stream they were captured on."""
        with torch.cuda.stream(self.stream):
            self.graph.replay()
        self.state.step_index += self.steps


__all__ = ["Gm1LargeStepGraph"]
