"""Ensembles of perturbed storms on uneven grids, advanced together on one GPU.

Each member is a complete large-step state (its own grid, fields, Kessler
and dry-mass workspaces).  A member's whole large step -- three RK stages
with all their acoustic small steps, the Kessler transition and solve3 --
is captured once into a HIP graph; a batched step replays every member's
graph, the members spread round-robin over a few streams so that small
members' kernels overlap.  The host then issues one graph launch per
member per step instead of several hundred kernel launches, which is what
bounds a sequential run of small members.

Every member's words are those of its own eager large step: a replay runs
the captured launches on the member's buffers, and members share no buffer.

Ordering: ``step`` forks once -- every member stream waits for the caller's
current stream and for each member state's bound stream -- then replays
all members, then joins once: the current stream and every member's bound
stream wait for all member streams.  A perturbation or upload queued on
either before ``step`` is seen by the replay, and work queued after it sees
the advanced members, with no host synchronisation.
"""

from __future__ import annotations

from typing import Any, Sequence

import torch

from .fields import CaseConfig
from .graph import Gm1LargeStepGraph
from .large_step import Gm1LargeStepState, create_gm1_large_step_state, host_initial_state, launch_cuda_gm1_large_step


class Gm1Ensemble:
    """Members' large steps as HIP graphs replayed over ``streams`` streams (default priority)."""

    def __init__(self, states: Sequence[Gm1LargeStepState], streams: int = 4, steps_per_graph: int = 1):
        if not states:
            raise ValueError("an ensemble needs at least one member")
        if streams < 1:
            raise ValueError("streams must be >= 1")
        self.states = list(states)
        self.streams = [torch.cuda.Stream() for _ in range(min(streams, len(self.states)))]
        self.graphs = [Gm1LargeStepGraph(st, steps=steps_per_graph, stream=self.streams[m % len(self.streams)])
                       for m, st in enumerate(self.states)]
        self.steps_per_graph = steps_per_graph

    @property
    def cells(self) -> int:
        return sum(st.cfg.cells for st in self.states)

    def _callers(self) -> list:
        cur = torch.cuda.current_stream()
        out = [cur]
        for st in self.states:
            if st.stream is not None and all(st.stream != s for s in out):
                out.append(st.stream)
        return [s for s in out if all(s != m for m in self.streams)]

    def step(self) -> None:
        """Enqueue ``steps_per_graph`` large steps of every member (asynchronous, fork/join ordered)."""
        callers = self._callers()
        for m in self.streams:
            for c in callers:
                m.wait_stream(c)
        for g in self.graphs:
            g.replay(ordered=False)
        for c in callers:
            for m in self.streams:
                c.wait_stream(m)

    def synchronize(self) -> None:
        for s in self.streams:
            s.synchronize()


def create_ensemble_states(members: Sequence[CaseConfig], device: Any = "cuda",
                           autotune: bool = False) -> list[Gm1LargeStepState]:
    """Initialise and upload every member (host fields built one member at a time).

    ``autotune`` picks each member's column pass by timing (``autotune.tune_column_pass``, once per grid
    shape); the variants give the same words."""
    out = []
    for cfg in members:
        host = host_initial_state(cfg)
        st = create_gm1_large_step_state(host, device)
        del host
        if autotune:
            from .autotune import autotune_state

            autotune_state(st)
        out.append(st)
    return out


def sequential_steps(states: Sequence[Gm1LargeStepState], steps: int) -> None:
    """The plain form: each member's eager large steps, one member after another on one stream."""
    for st in states:
        for _ in range(steps):
            launch_cuda_gm1_large_step(st)


__all__ = ["Gm1Ensemble", "create_ensemble_states", "sequential_steps"]
