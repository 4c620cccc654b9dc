"""BatchedWorld: thousands of independent worlds stepped by one launch sequence.

All worlds of the batch share dt and the iteration counts; each keeps its own
gravity, sleep and warm-starting flags.  The state stays on the GPU between
steps; ``state()`` copies it back in the batch layout.

Steps are graph-captured by default.  While the launch tuner is on, it takes
one eager step per candidate launch shape at the start of each tuning period
(real steps: the shape changes the speed, never the results).  After every
call the device's error flags are checked: a capacity error re-creates the
batch with more contact slots and replays the steps from the last good state
(``Stepper``); a non-finite body state raises ``B2GError`` (B2G_E_NAN).
"""
from __future__ import annotations

from typing import List

import numpy as np

from . import layout as L
from .autotune import LaunchTuner
from .builder import scenes_batch
from .stepper import Stepper


class BatchedWorld(object):
    def __init__(self, batch: L.BatchArrays, precision: str = "fp32", device: int = 0,
                 contact_factor: float = 8.0, contact_extra: int = 64, autotune: bool = True):
        self.batch = batch.normalized()
        self.precision = precision
        self.stepper = Stepper(precision, device, contact_factor, contact_extra)
        self.stepper.upload(self.batch)
        self.tuner = LaunchTuner(self.stepper.engine) if autotune else None
        self.stepper.on_regrow = self._regrown
        self.steps = 0

    def _regrown(self, engine) -> None:
        if self.tuner is not None:
            self.tuner.engine = engine
            if self.tuner.best:
                self.tuner.apply(self.tuner.best)

    @classmethod
    def from_scenes(cls, scenes: List[dict], precision: str = "fp32", device: int = 0, **kw) -> "BatchedWorld":
        return cls(scenes_batch(scenes), precision, device, **kw)

    @property
    def engine(self):
        return self.stepper.engine

    @property
    def n_worlds(self) -> int:
        return self.batch.n_worlds

    @property
    def dynamic_bodies(self) -> int:
        return int(np.count_nonzero(self.batch.body_i[:, L.BI_TYPE] == L.DYNAMIC))

    def set_param(self, name: str, value: int) -> None:
        """A launch-shape or debugging parameter of the engine (Engine.set_param)."""
        self.stepper.set_param(name, value)

    def step(self, dt: float = 1.0 / 60, vel_iters: int = 8, pos_iters: int = 3, n: int = 1,
             graph: bool = True) -> None:
        """n steps of every world (World.step(dt, vel_iters, pos_iters) each)."""
        if self.tuner is not None and graph:
            self.tuner.step_n(n, dt, vel_iters, pos_iters)
            self.stepper.record(n, dt, vel_iters, pos_iters)
            self.stepper.check()
        else:
            self.stepper.step(dt, vel_iters, pos_iters, n, graph)
        self.steps += n

    def query_aabb(self, worlds, boxes):
        """World.query_aabb for many boxes at once: rows (query, world-local fixture) of every fixture
        whose fat AABB overlaps box k = (lower x, lower y, upper x, upper y) of world worlds[k]."""
        return self.stepper.engine.query_aabb(worlds, boxes)

    def ray_cast(self, worlds, segments, closest: bool = True):
        """World.ray_cast for many segments (p1 x, p1 y, p2 x, p2 y) at once: each ray's closest hit, or
        every hit (Engine.ray_cast)."""
        return self.stepper.engine.ray_cast(worlds, segments, closest)

    def sync(self) -> None:
        self.stepper.check()

    def state(self) -> L.BatchArrays:
        return self.stepper.download()

    def close(self) -> None:
        self.stepper.close()
