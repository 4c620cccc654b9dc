"""Launch-shape autotuner for the island solver.

The island solver's results do not depend on its launch shape
(``Engine.set_param``: which islands get a team of lanes, how large the teams
are, how much LDS a wavefront reserves), only its speed does, and the best
shape depends on the island sizes of the current state: the reserved LDS sets
how many wavefronts fit on a compute unit, and a team whose island does not
fit its share falls back to one lane.  The tuner therefore runs real
simulation steps, one per candidate shape, times each step's solve kernel
with HIP events, and keeps the fastest.  No step is wasted: a tuning step
advances the simulation like any other step.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence

import numpy as np

CANDIDATES: Sequence[Dict[str, int]] = (
    {"heavy_class": 5, "team_full_class": 9, "team_max": 32, "team_lds": 24576},
    {"heavy_class": 5, "team_full_class": 9, "team_max": 32, "team_lds": 20480},
    {"heavy_class": 5, "team_full_class": 9, "team_max": 32, "team_lds": 32768},
    {"heavy_class": 5, "team_full_class": 9, "team_max": 64, "team_lds": 20480},
    {"heavy_class": 5, "team_full_class": 8, "team_max": 64, "team_lds": 20480},
    {"heavy_class": 5, "team_full_class": 10, "team_max": 32, "team_lds": 24576},
    {"heavy_class": 5, "team_full_class": 7, "team_max": 32, "team_lds": 20480},
)


class LaunchTuner(object):
    """Re-tunes the solver launch shape during `step_n`: after `early` steps at first (while a
    scene is still changing fast), then every `period` steps."""

    def __init__(self, engine, candidates: Sequence[Dict[str, int]] = CANDIDATES, period: int = 200,
                 early: Sequence[int] = (50, 50, 100)):
        self.engine = engine
        self.candidates: List[Dict[str, int]] = [dict(c) for c in candidates]
        self.period = max(period, len(self.candidates) + 1)
        self.early = [max(e, len(self.candidates) + 1) for e in early]
        self.tunes = 0
        self.best: Optional[Dict[str, int]] = None
        self.last_times: Optional[np.ndarray] = None
        self.last_ms: Optional[np.ndarray] = None  # per candidate: collide, islands, solve, sync, pairs
        self._since = None

    def interval(self) -> int:
        """Steps from the last tuning (its own steps included) to the next one."""
        k = self.tunes - 1
        return self.early[k] if 0 <= k < len(self.early) else self.period

    def apply(self, params: Dict[str, int]) -> None:
        for k, v in params.items():
            self.engine.set_param(k, v)

    def tune(self, dt: float, vel_iters: int, pos_iters: int) -> Dict[str, int]:
        """Steps once per candidate (eagerly, timed) and applies the fastest."""
        ms = np.zeros((len(self.candidates), 5))
        for i, c in enumerate(self.candidates):
            self.apply(c)
            ms[i] = self.engine.profile(1, dt, vel_iters, pos_iters)
        times = ms[:, 2]  # the solve kernel
        self.last_ms = ms
        self.last_times = times
        self.best = self.candidates[int(np.argmin(times))]
        self.apply(self.best)
        self._since = 0
        self.tunes += 1
        return self.best

    def step_n(self, n: int, dt: float = 1.0 / 60, vel_iters: int = 8, pos_iters: int = 3) -> None:
        """n steps: tuning steps at the start of every period, graph-captured steps otherwise."""
        done = 0
        while done < n:
            if self._since is None or self._since >= self.interval():
                if n - done < len(self.candidates):
                    self.engine.step_n(n - done, dt, vel_iters, pos_iters, graph=True)
                    self._since = (self._since or 0) + n - done
                    return
                self.tune(dt, vel_iters, pos_iters)
                done += len(self.candidates)
                self._since = len(self.candidates)
                continue
            k = min(n - done, self.interval() - self._since)
            self.engine.step_n(k, dt, vel_iters, pos_iters, graph=True)
            done += k
            self._since += k
