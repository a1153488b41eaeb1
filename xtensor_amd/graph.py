"""A sequence of prepared plans replayed as one HIP graph.

`PlanGraph([plan(c, a * b + exp(-a * a)), plan(s, sum(c, [1])), ...])`
captures the plans' launches once and `replay()` enqueues all of them with
one call on the current stream.  It is for sequences of small launches whose
cost is dominated by launch overhead; the kernels and their results are those
of the plans themselves.

The plans keep their operands alive and read whatever those hold at replay
time, as `Plan.run()` does.  Before capturing, every plan runs once on the
graph's own stream, so one-time work (run-time compiles, workspace growth)
stays out of the graph.  The captured launches use that stream's workspace:
replays of one `PlanGraph` must not overlap (successive replays on one stream
are fine).
"""

from __future__ import annotations

from typing import Sequence

from .core import Plan


class PlanGraph:
    def __init__(self, plans: Sequence[Plan]):
        import torch

        if not plans:
            raise ValueError("PlanGraph needs at least one plan")
        self.plans = list(plans)
        self._stream = torch.cuda.Stream()  # default priority
        self._stream.wait_stream(torch.cuda.current_stream())  # operands written so far
        handle = self._stream.cuda_stream
        # one eager pass: run-time compiles and workspace growth happen here
        for p in self.plans:
            p.run(stream=handle)
        self._stream.synchronize()
        self._graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self._graph, stream=self._stream):
            for p in self.plans:
                p.run(stream=handle)
        torch.cuda.current_stream().wait_stream(self._stream)

    def replay(self) -> None:
        """Enqueue every captured launch, in order, on the current stream."""
        self._graph.replay()
