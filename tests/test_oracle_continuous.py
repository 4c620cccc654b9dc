"""The oracle's continuous physics (World.solve_toi) against recorded reference runs and closed forms.

Fixtures (tests/golden/make_golden.py, pypybox2d's pure-Python flavour, with
continuous_physics on, the reference's default):
  * continuous_projectiles_150.npz: fast bodies, every other one a bullet, shot
    at a thin static wall and into the ground, a column of boxes and a
    kinematic paddle; 150 steps;
  * continuous_pyramid_200.npz: the 10-row pyramid, whose boxes land on the
    ground in the first steps; 200 steps.
Both record each step's new contacts in the reference's creation order, the
states at checkpoints, and per step the number of time_of_impact calls and of
TOI events (sub-step island solves).  The oracle replays them with the
recorded creation order and must match every checkpoint state bit for bit and
every step's counts exactly.  The closed forms pin the time of impact itself.
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as O
from pypybox2d_amd import builder, layout as L

RUNS = ["continuous_projectiles_150.npz", "continuous_pyramid_200.npz"]


@pytest.mark.parametrize("run", RUNS)
def test_oracle_replays_continuous_reference_runs_bit_exact(run):
    z = np.load(os.path.join(GOLDEN, run))
    init = L.BatchArrays(*[z["init_" + n] for n in L._FIELDS])
    assert init.world_i[0, L.WI_FLAGS] & L.WF_CONTINUOUS
    o = O.Oracle(init)
    hoff, hp = z["hint_off"], z["hint_pairs"]
    cks = set(int(s) for s in z["checkpoints"])
    events = 0
    for s in range(1, int(z["meta_steps"]) + 1):
        o.set_hint(0, hp[hoff[s - 1]:hoff[s]])
        assert o.step() == 0, s
        assert o.toi_stats(0) == tuple(int(v) for v in z["toi_stats"][s - 1]), s
        events += o.toi_stats(0)[1]
        if s in cks:
            st = o.state()
            for name in ("body_d", "fix_d", "con_d", "con_i", "move"):
                got, want = getattr(st, name), z["s%d_%s" % (s, name)]
                assert got.shape == want.shape and np.array_equal(got, want), (s, name)
    assert events > 10


def _scene(bodies):
    return {"gravity": (0.0, 0.0), "do_sleep": False, "warm_starting": True, "bodies": bodies}


def _body(type_, pos, shape, vel=(0.0, 0.0), bullet=False):
    b = {"type": type_, "position": pos, "angle": 0.0, "linear_velocity": vel, "angular_velocity": 0.0,
         "fixtures": [{"shape": shape, "density": 1.0 if type_ == "dynamic" else 0.0, "friction": 0.2,
                       "restitution": 0.0}]}
    if bullet:
        b["bullet"] = True
    return b


def _continuous(scene):
    b = builder.scene_batch(scene)
    b.world_i[0, L.WI_FLAGS] |= L.WF_CONTINUOUS
    return b


def test_a_fast_circle_stops_at_a_thin_wall_instead_of_tunnelling():
    """A circle of radius 0.5 moving at 120 m/s (2 m per step) toward a static wall 0.2 m thick
    (faces at x = 9.9 and 10.1).  Discrete steps carry it through: after step 1 it is at 8.6, short
    of the wall, after step 2 at 10.6, past its middle, and the contact pushes it on.  With
    continuous physics its contact (created at the end of step 1, tested from step 2 on) gives a
    time of impact in step 2: the circle is advanced to where its centre is the TOI target from the
    wall's core, max(slop, r + skin - 3 slop) = 0.495 (toi.py:262-265), then sub-stepped, which
    leaves it against the wall's skin (9.9 - 0.5 - 0.01, less the 0.005 slop the position solver
    allows) with its normal velocity removed."""
    scene = _scene([_body("static", (10.0, 0.0), ("box", 0.1, 2.0)),
                    _body("dynamic", (6.6, 0.0), ("circle", 0.5), vel=(120.0, 0.0))])
    discrete = O.Oracle(builder.scene_batch(scene))
    discrete.step()
    discrete.step()
    assert discrete.state().body_d[1, L.BD_C_X] > 10.5  # through the wall's middle
    o = O.Oracle(_continuous(scene))
    o.step()
    assert o.toi_stats(0) == (0, 0)  # the new contact is not enabled before its first update
    o.step()
    calls, events = o.toi_stats(0)
    assert calls >= 1 and events == 1
    st = o.state()
    x = st.body_d[1, L.BD_C_X]
    assert 9.9 - 0.5 - 0.01 - 0.0051 <= x <= 9.9 - 0.495 + 0.25 * 0.005, x
    assert abs(st.body_d[1, L.BD_V_X]) < 1.0
    for _ in range(30):
        o.step()
    assert abs(o.state().body_d[1, L.BD_C_X] - x) < 1e-3


def test_slow_and_non_bullet_pairs_need_no_time_of_impact():
    """Only contacts with a bullet, static or kinematic body are tested (world.py:647-653): a fast
    circle and a dynamic slab (contact created in step 1, overlapping in step 2) get no time of
    impact, unless the circle is a bullet."""
    scene = _scene([_body("dynamic", (0.0, 0.0), ("box", 0.1, 2.0)),
                    _body("dynamic", (-4.2, 0.0), ("circle", 0.5), vel=(120.0, 0.0))])
    o = O.Oracle(_continuous(scene))
    for _ in range(2):
        o.step()
        assert o.toi_stats(0) == (0, 0)
    scene["bodies"][1]["bullet"] = True
    o = O.Oracle(_continuous(scene))
    o.step()
    o.step()
    assert o.toi_stats(0)[1] == 1


def test_a_resting_contact_is_tested_every_step_without_events():
    """A box resting on a static ground: the contact is tested every step (a static body is
    involved), and once the box rests on its skin (core 0.015 m above the edge's core: the 0.02 m
    of the two skins less the 0.005 m slop) the sweep stays separated: no TOI event."""
    scene = {"gravity": (0.0, -10.0), "do_sleep": False, "warm_starting": True,
             "bodies": [_body("static", (0.0, 0.0), ("edge", (-5.0, 0.0), (5.0, 0.0))),
                        _body("dynamic", (0.0, 0.515), ("box", 0.5, 0.5))]}
    o = O.Oracle(_continuous(scene))
    o.step()
    for s in range(2, 121):
        o.step()
        calls, events = o.toi_stats(0)
        assert calls == 1 and events == 0, (s, calls, events)
    assert math.isclose(o.state().body_d[1, L.BD_C_Y], 0.515, abs_tol=1e-3)
