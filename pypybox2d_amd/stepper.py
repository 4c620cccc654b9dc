"""A device batch that grows its contact capacity when a step runs out of it.

Each world of a batch has a fixed number of contact slots on the device
(``contact_factor`` per fixture plus ``contact_extra``).  When a step needs
more (more fat-AABB pairs than slots, or more new pairs than the pair search
holds), the kernels flag the world and end normally, and the next sync reports
``B2G_E_CAPACITY``; that step's result is not usable.

``Stepper`` keeps the last state that is known to be good on the host (the
uploaded batch, or the last download) and the operations run since then:
steps, pair searches run ahead of a step, and drops of new contacts (a custom
contact filter's verdicts).  On a capacity error it re-creates the device
batch with twice the capacity, uploads that state and replays them.  The
pipeline is deterministic and its results do not depend on the capacity, the
launch shape or graph capture, so the replay gives exactly what the step would
have given with enough room.
Other errors (non-finite state, unsupported shapes) are raised.

The reference has no capacity limits (Python lists grow); this is the
device-side counterpart of that growth.
"""
from __future__ import annotations

from typing import Dict, List, Optional

from . import layout as L
from ._native import B2G_E_CAPACITY, B2GError, Engine

MAX_REGROW = 6
MAX_LOG = 256  # operations kept for a replay before the state is checkpointed by a download


class Stepper(object):
    def __init__(self, precision: str = "fp32", device: int = 0, contact_factor: float = 8.0,
                 contact_extra: int = 64, engine_factory=Engine):
        self.precision, self.device = precision, device
        self.contact_factor, self.contact_extra = float(contact_factor), int(contact_extra)
        self._new_engine = engine_factory
        self.engine = engine_factory(precision, device, contact_factor=self.contact_factor,
                                     contact_extra=self.contact_extra)
        self.params: Dict[str, int] = {}
        self._ckpt: Optional[L.BatchArrays] = None
        self._log: List[tuple] = []  # operations since _ckpt: ("step", n, dt, vi, pi), ("pairs",), ("drop", w, j)
        self.regrows = 0
        self.on_regrow = None  # called with the new Engine after a regrow
        self._split = False     # a step begun by collide() is waiting for solve()

    def set_param(self, name: str, value: int) -> None:
        self.engine.set_param(name, value)
        if name not in ("island_times",):
            self.params[name] = int(value)

    def upload(self, batch: L.BatchArrays) -> None:
        self.engine.upload(batch)
        self._ckpt = batch.normalized().copy()
        self._log = []

    def download(self) -> L.BatchArrays:
        st = self.engine.download()
        if not self._split:  # a state between a step's two parts is no replay start
            self._ckpt = st.copy()
            self._log = []
        return st

    def record(self, n: int, dt: float, vel_iters: int, pos_iters: int) -> None:
        """Logs n steps that were launched on the engine directly (e.g. by the launch tuner)."""
        last = self._log[-1] if self._log else None
        if last is not None and last[0] == "step" and last[2:] == (dt, vel_iters, pos_iters):
            self._log[-1] = ("step", last[1] + n, dt, vel_iters, pos_iters)
        else:
            self._log.append(("step", n, dt, vel_iters, pos_iters))

    def pair_search(self) -> None:
        """The pending new-fixture pair search, ahead of the next step (Engine.pair_search)."""
        self.engine.pair_search()
        self._log.append(("pairs",))
        self.check()

    def new_contacts(self):
        return self.engine.new_contacts()

    def drop(self, worlds, index) -> None:
        """Drops new contacts of the last pair search (Engine.drop_contacts)."""
        self.engine.drop_contacts(worlds, index)
        self._log.append(("drop", list(worlds), list(index)))

    def events(self):
        return self.engine.events()

    def collide(self, dt: float, vel_iters: int, pos_iters: int) -> None:
        """The first part of a step (Engine.step_collide): up to and with the narrow phase."""
        self.engine.step_collide(dt, vel_iters, pos_iters)
        self._split = True
        self._log.append(("collide", dt, vel_iters, pos_iters))
        self.check()

    def disable(self, worlds, slots) -> None:
        """Contacts a pre_solve disabled for this step (Engine.disable_contacts)."""
        self.engine.disable_contacts(worlds, slots)
        self._log.append(("disable", list(worlds), list(slots)))

    def solve(self) -> None:
        """The rest of the step begun by collide()."""
        self.engine.step_solve()
        self._split = False
        self._log.append(("solve",))
        self.check()

    def _replay(self, op) -> None:
        if op[0] == "step":
            self.engine.step_n(op[1], op[2], op[3], op[4], True)
        elif op[0] == "pairs":
            self.engine.pair_search()
        elif op[0] == "collide":
            self.engine.step_collide(op[1], op[2], op[3])
        elif op[0] == "disable":
            self.engine.disable_contacts(op[1], op[2])
        elif op[0] == "solve":
            self.engine.step_solve()
        else:
            self.engine.sync()  # the drop's indices refer to the search just replayed
            self.engine.drop_contacts(op[1], op[2])

    def step(self, dt: float, vel_iters: int, pos_iters: int, n: int = 1, graph: bool = False) -> None:
        """n steps, then a check of the device's error flags (regrowing on a capacity error)."""
        if n == 1 and not graph:
            self.engine.step(dt, vel_iters, pos_iters)
        else:
            self.engine.step_n(n, dt, vel_iters, pos_iters, graph)
        self.record(n, dt, vel_iters, pos_iters)
        self.check()

    def check(self) -> None:
        try:
            self.engine.sync()
        except B2GError as e:
            if e.code != B2G_E_CAPACITY:
                raise
            self._regrow(e)
        if len(self._log) > MAX_LOG and not self._split:
            self.download()

    def _regrow(self, first: B2GError) -> None:
        if self._ckpt is None:
            raise first
        err = first
        for _ in range(MAX_REGROW):
            self.contact_factor *= 2.0
            self.contact_extra *= 2
            self.engine.close()
            self.engine = self._new_engine(self.precision, self.device, contact_factor=self.contact_factor,
                                           contact_extra=self.contact_extra)
            for k, v in self.params.items():
                self.engine.set_param(k, v)
            self.engine.upload(self._ckpt)
            self.regrows += 1
            try:
                for op in self._log:
                    self._replay(op)
                self.engine.sync()
            except B2GError as e:
                if e.code != B2G_E_CAPACITY:
                    raise
                err = e
                continue
            if self.on_regrow is not None:
                self.on_regrow(self.engine)
            return
        raise err

    def close(self) -> None:
        self.engine.close()
