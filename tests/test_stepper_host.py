"""Host logic of the capacity regrow-and-replay (pypybox2d_amd/stepper.py), with a stand-in engine
on the CPU: a step that needs more contact slots than the engine has is reported at the next sync,
and the stepper re-creates the engine with twice the slots, uploads the last good state and replays
exactly the logged steps."""
import numpy as np
import pytest

from pypybox2d_amd import layout as L
from pypybox2d_amd._native import B2G_E_CAPACITY, B2G_E_NAN, B2GError
from pypybox2d_amd.stepper import Stepper


class FakeEngine:
    """Needs `need` slots per fixture from step `at` on; records what it was asked to do."""
    made = []

    def __init__(self, precision, device, contact_factor=8.0, contact_extra=64, need=10.0, at=3, nan_at=None):
        self.factor, self.calls, self.params, self.steps = contact_factor, [], {}, 0
        self.need, self.at, self.nan_at = need, at, nan_at
        FakeEngine.made.append(self)

    def upload(self, b):
        self.calls.append(("upload", int(b.body_d[0, 0])))
        self.steps = int(b.body_d[0, 0])

    def step(self, dt, vi, pi):
        self.step_n(1, dt, vi, pi, False)

    def step_n(self, n, dt, vi, pi, graph):
        self.calls.append(("step", n, dt, vi, pi))
        self.steps += n

    def sync(self):
        if self.nan_at is not None and self.steps >= self.nan_at:
            raise B2GError(B2G_E_NAN, "nan")
        if self.steps >= self.at and self.factor < self.need:
            raise B2GError(B2G_E_CAPACITY, "capacity")

    def set_param(self, k, v):
        self.params[k] = v

    def pair_search(self):
        self.calls.append(("pairs",))

    def drop_contacts(self, w, j):
        self.calls.append(("drop", list(w), list(j)))

    def download(self):
        b = start_state()
        b.body_d[0, 0] = self.steps  # the "state" is the number of steps taken
        return b

    def close(self):
        pass


def start_state():
    b = L.BatchArrays.empty_world()
    b.body_d = np.zeros((1, L.BODY_ND))
    b.body_off = np.array([0, 1], np.int32)
    b.body_i = np.zeros((1, L.BODY_NI), np.int32)
    return b


def test_capacity_error_regrows_and_replays_from_the_last_good_state():
    FakeEngine.made = []
    s = Stepper("fp32", 0, contact_factor=2.0, contact_extra=8, engine_factory=FakeEngine)
    s.set_param("team_max", 32)
    s.upload(start_state())
    s.step(0.01, 8, 3)
    s.step(0.01, 8, 3, n=4, graph=True)  # crosses step 3 -> capacity error -> regrow to 4, 8, 16
    assert s.regrows == 3 and s.contact_factor == 16.0 and s.contact_extra == 64
    last = FakeEngine.made[-1]
    assert last.factor == 16.0 and last.params == {"team_max": 32}
    # the replay: the start state, then the logged steps with their dt and iteration counts
    assert last.calls == [("upload", 0), ("step", 5, 0.01, 8, 3)]
    assert s.download().body_d[0, 0] == 5
    # a new checkpoint: later replays start from the downloaded state
    s.step(0.02, 6, 2)
    assert s._log == [("step", 1, 0.02, 6, 2)] and int(s._ckpt.body_d[0, 0]) == 5


def test_other_errors_are_raised_and_not_retried():
    FakeEngine.made = []

    def factory(*a, **k):
        return FakeEngine(*a, nan_at=2, **k)
    s = Stepper("fp32", 0, engine_factory=factory)
    s.upload(start_state())
    s.step(0.01, 8, 3)
    with pytest.raises(B2GError) as e:
        s.step(0.01, 8, 3)
    assert e.value.code == B2G_E_NAN and s.regrows == 0 and len(FakeEngine.made) == 1


def test_regrow_gives_up_after_a_bounded_number_of_attempts():
    def factory(*a, **k):
        return FakeEngine(*a, need=1e12, at=1, **k)
    s = Stepper("fp32", 0, engine_factory=factory)
    s.upload(start_state())
    with pytest.raises(B2GError) as e:
        s.step(0.01, 8, 3)
    assert e.value.code == B2G_E_CAPACITY


def test_replay_repeats_pair_searches_and_drops_in_order():
    """A custom filter's work between steps (a pair search run ahead of a step, drops of new
    contacts) is part of the replay, in the order it happened."""
    FakeEngine.made = []
    s = Stepper("fp32", 0, contact_factor=4.0, engine_factory=lambda *a, **k: FakeEngine(*a, need=8.0, at=3, **k))
    s.upload(start_state())
    s.pair_search()
    s.drop([0, 0], [1, 3])
    s.step(0.01, 8, 3)
    s.drop([0], [0])
    s.step(0.01, 8, 3, n=2)  # reaches step 3 -> regrow to factor 8
    assert s.regrows == 1
    assert FakeEngine.made[-1].calls == [("upload", 0), ("pairs",), ("drop", [0, 0], [1, 3]), ("step", 1, 0.01, 8, 3),
                                         ("drop", [0], [0]), ("step", 2, 0.01, 8, 3)]


def test_a_long_log_is_checkpointed():
    from pypybox2d_amd import stepper
    FakeEngine.made = []
    s = Stepper("fp32", 0, engine_factory=FakeEngine)
    s.upload(start_state())
    for k in range(stepper.MAX_LOG + 1):
        s.drop([0], [0])
        s.step(0.01, 8, 3)
    assert len(s._log) <= stepper.MAX_LOG and int(s._ckpt.body_d[0, 0]) > 0
