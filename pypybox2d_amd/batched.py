"""BatchedWorld: thousands of independent worlds stepped by one launch sequence.

All worlds of the batch share dt and the iteration counts; each keeps its own
gravity, sleep and warm-starting flags.  The state stays on the GPU between
steps; ``state()`` copies it back in the batch layout.
"""
from __future__ import annotations

from typing import List, Optional

import numpy as np

from . import layout as L
from ._native import Engine
from .autotune import LaunchTuner
from .builder import scenes_batch


class BatchedWorld(object):
    def __init__(self, batch: L.BatchArrays, precision: str = "fp32", device: int = 0,
                 contact_factor: float = 8.0, autotune: bool = True):
        self.batch = batch.normalized()
        self.engine = Engine(precision, device, contact_factor=contact_factor)
        self.engine.upload(self.batch)
        self.precision = precision
        # the solver's launch shape changes its speed, never its results
        self.tuner = LaunchTuner(self.engine) if autotune else None

    @classmethod
    def from_scenes(cls, scenes: List[dict], precision: str = "fp32", device: int = 0) -> "BatchedWorld":
        return cls(scenes_batch(scenes), precision, device)

    @property
    def n_worlds(self) -> int:
        return self.batch.n_worlds

    @property
    def dynamic_bodies(self) -> int:
        return int(np.count_nonzero(self.batch.body_i[:, L.BI_TYPE] == L.DYNAMIC))

    def step(self, dt: float = 1.0 / 60, vel_iters: int = 8, pos_iters: int = 3, n: int = 1,
             graph: bool = True) -> None:
        if n == 1 and not graph:
            self.engine.step(dt, vel_iters, pos_iters)
        elif self.tuner is not None and graph:
            self.tuner.step_n(n, dt, vel_iters, pos_iters)
        else:
            self.engine.step_n(n, dt, vel_iters, pos_iters, graph)

    def sync(self) -> None:
        self.engine.sync()

    def state(self) -> L.BatchArrays:
        return self.engine.download()

    def close(self) -> None:
        self.engine.close()
