"""Launch-shape autotuner for the step's kernels.

No launch shape changes a result (the worlds and islands are independent and
the order of work inside each is fixed); only the speed depends on them, and
the best shapes depend on the current state: island sizes, how many contacts
and new pairs the worlds hold.  The tuner therefore runs real simulation
steps (no step is wasted: a tuning step advances the simulation like any
other), times each stage with HIP events (``Engine.profile``) and keeps, for
each group of parameters, the candidate whose stage was fastest:

  solve     which islands get a team of lanes, how large the teams are and
            how much LDS a solver wavefront reserves (the island solve);
  collide   contact slots per narrow-phase workgroup (the pair-buffer tile
            of the narrow phase);
  islands   which island-search launches use 256- rather than 128-thread
            workgroups;
  pairs     which pair-search launches use 1024- rather than 256-thread
            workgroups.

Each tuning pass takes one step per solve candidate; step i also applies the
(i mod n)-th candidate of every other group, and a group's choice is the
candidate with the smallest mean time of its stage over the steps it ran in.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence

import numpy as np

SOLVE_CANDIDATES: Sequence[Dict[str, int]] = (
    {"heavy_class": 5, "team_full_class": 9, "team_max": 32, "team_lds": 24576},
    {"heavy_class": 5, "team_full_class": 9, "team_max": 32, "team_lds": 20480},
    {"heavy_class": 5, "team_full_class": 9, "team_max": 32, "team_lds": 32768},
    {"heavy_class": 5, "team_full_class": 9, "team_max": 64, "team_lds": 20480},
    {"heavy_class": 5, "team_full_class": 8, "team_max": 64, "team_lds": 20480},
    {"heavy_class": 5, "team_full_class": 10, "team_max": 32, "team_lds": 24576},
    {"heavy_class": 5, "team_full_class": 7, "team_max": 32, "team_lds": 20480},
)
CANDIDATES = SOLVE_CANDIDATES

# (group, stage index of Engine.profile, candidates)
GROUPS = (
    ("solve", 2, SOLVE_CANDIDATES),
    ("collide", 0, ({"collide_block": 256}, {"collide_block": 128}, {"collide_block": 512})),
    ("islands", 1, ({"islands_wide": 5}, {"islands_wide": 0}, {"islands_wide": 32})),
    ("pairs", 4, ({"pairs_wide": 2}, {"pairs_wide": 0}, {"pairs_wide": 4})),
)


class LaunchTuner(object):
    """Re-tunes the launch shapes during `step_n`: after `early` steps at first (while a
    scene is still changing fast), then every `period` steps."""

    def __init__(self, engine, groups=GROUPS, period: int = 200, early: Sequence[int] = (50, 50, 100)):
        self.engine = engine
        self.groups = [(name, stage, [dict(c) for c in cands]) for name, stage, cands in groups]
        self.steps_per_tune = max(len(c) for _, _, c in self.groups)
        self.period = max(period, self.steps_per_tune + 1)
        self.early = [max(e, self.steps_per_tune + 1) for e in early]
        self.tunes = 0
        self.best: Optional[Dict[str, int]] = None
        self.choice: Dict[str, Dict[str, int]] = {}  # per group
        self.last_times: Optional[np.ndarray] = None  # solve ms per solve candidate
        self.last_ms: Optional[np.ndarray] = None  # per tuning step: the five stage times
        self._since = None

    @property
    def candidates(self) -> List[Dict[str, int]]:
        """The solve candidates (one tuning step each)."""
        return self.groups[0][2]

    def interval(self) -> int:
        """Steps from the last tuning (its own steps included) to the next one."""
        k = self.tunes - 1
        return self.early[k] if 0 <= k < len(self.early) else self.period

    def apply(self, params: Dict[str, int]) -> None:
        for k, v in params.items():
            self.engine.set_param(k, v)

    def tune(self, dt: float, vel_iters: int, pos_iters: int) -> Dict[str, int]:
        """Steps once per solve candidate (eagerly, timed) and applies each group's fastest."""
        n = self.steps_per_tune
        ms = np.zeros((n, 5))
        for i in range(n):
            for _, _, cands in self.groups:
                self.apply(cands[i % len(cands)])
            ms[i] = self.engine.profile(1, dt, vel_iters, pos_iters)
        self.last_ms = ms
        best: Dict[str, int] = {}
        for name, stage, cands in self.groups:
            t = np.array([ms[[i for i in range(n) if i % len(cands) == c], stage].mean() for c in range(len(cands))])
            if name == "solve":
                self.last_times = t
            self.choice[name] = cands[int(np.argmin(t))]
            best.update(self.choice[name])
        self.best = best
        self.apply(best)
        self._since = 0
        self.tunes += 1
        return best

    def step_n(self, n: int, dt: float = 1.0 / 60, vel_iters: int = 8, pos_iters: int = 3) -> None:
        """n steps: tuning steps at the start of every period, graph-captured steps otherwise."""
        done = 0
        while done < n:
            if self._since is None or self._since >= self.interval():
                if n - done < self.steps_per_tune:
                    self.engine.step_n(n - done, dt, vel_iters, pos_iters, graph=True)
                    self._since = (self._since or 0) + n - done
                    return
                self.tune(dt, vel_iters, pos_iters)
                done += self.steps_per_tune
                self._since = self.steps_per_tune
                continue
            k = min(n - done, self.interval() - self._since)
            self.engine.step_n(k, dt, vel_iters, pos_iters, graph=True)
            done += k
            self._since += k
