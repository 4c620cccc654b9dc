"""GPU parity of continuous physics (World.solve_toi after each step, k_toi) against the oracle.

The oracle reproduces the reference's continuous-physics runs bit for bit
(test_oracle_continuous.py).  Here the device runs the same scenes from the
same state and is held to the oracle step by step:
  * every step's time_of_impact calls and TOI events (b2g_stats) equal the
    oracle's, and the contact lists are the same at the checkpoints;
  * FP64 body states within 1e-9 (relative, floors as in parity_util) for the
    first 60 steps, 1e-6 after (the device's sin/cos differ from libm's in the
    last bit now and then, and impacts amplify it);
  * fast bodies thrown at a thin wall stop or pass as in the oracle, in the
    FP32 pipeline as in FP64;
  * the drop-in World with continuous physics on (the reference's default)
    runs the reference's example scene, and its listener receives the calls of
    the time-of-impact sub-steps as the oracle records them.
"""
import numpy as np
import pytest

import parity_util as PU
from oracle import oracle as O
from pypybox2d_amd import _native as N
from pypybox2d_amd import builder, layout as L, scenes

import pypybox2d_amd as b2

pytestmark = pytest.mark.gpu

DT = 1.0 / 60


def continuous_batch(scene):
    b = builder.scene_batch(scene)
    b.world_i[0, L.WI_FLAGS] |= L.WF_CONTINUOUS
    return b


@pytest.mark.parametrize("name,steps", [("projectiles", 150), ("pyramid", 200)])
def test_continuous_runs_track_the_oracle_fp64(name, steps):
    batch = continuous_batch(getattr(scenes, name)())
    o = O.Oracle(batch)
    e = N.Engine("fp64")
    e.upload(batch)
    e.stats(reset=True)
    events = 0
    worst = 0.0
    for s in range(1, steps + 1):
        assert o.step(DT, 8, 3) == 0
        e.step(DT, 8, 3)
        st = e.stats(reset=True)
        want = o.toi_stats(0)
        assert (st["toi_calls"], st["toi_events"]) == want, (s, st["toi_calls"], st["toi_events"], want)
        events += want[1]
        if s % 10 == 0 or s == steps:
            ref, got = o.state(), e.download()
            assert PU.same_contact_lists(got, ref), s
            be = PU.body_errors(got, ref)
            err = max(be[q] for q in PU.QUANTITIES)
            worst = max(worst, err)
            assert err <= (1e-9 if s <= 60 else 1e-6), (s, be)
    print(name, "toi events", events, "worst body error", worst)
    assert events > 5
    e.close()


def wall_scene():
    """Eight small bodies (circles, boxes; a quarter of them bullets) flying at a thin static wall
    (0.2 m thick, faces at x = 9.9 and 10.1) at 80-150 m/s, without gravity."""
    bodies = [{"type": "static", "position": (10.0, 4.0), "angle": 0.0, "linear_velocity": (0.0, 0.0),
               "angular_velocity": 0.0, "fixtures": [{"shape": ("box", 0.1, 4.0), "density": 0.0,
                                                      "friction": 0.2, "restitution": 0.0}]}]
    for k in range(8):
        shape = ("circle", 0.15 + 0.05 * (k % 3)) if k % 2 == 0 else ("box", 0.15, 0.1 + 0.05 * (k % 3))
        bodies.append({"type": "dynamic", "position": (-3.0 - 1.7 * k, 0.5 + 0.9 * k), "angle": 0.3 * k,
                       "linear_velocity": (80.0 + 10.0 * k, 0.0), "angular_velocity": 2.0 - 0.5 * k,
                       "bullet": k % 4 == 0,
                       "fixtures": [{"shape": shape, "density": 1.0, "friction": 0.3, "restitution": 0.0}]})
    return {"gravity": (0.0, 0.0), "do_sleep": True, "warm_starting": True, "bodies": bodies}


def test_fast_bodies_meet_the_wall_as_in_the_oracle_in_both_pipelines():
    """Without continuous physics every body passes the wall.  With it, the bodies whose contact
    with the wall exists before the step that would carry them through are stopped by a time of
    impact; the others still pass, because a proxy's fat AABB only grows when its swept AABB leaves
    it (dynamictree.py:152-188), so the contact can come a step late, as in the reference.  The
    device stops exactly the bodies the oracle stops, in FP32 as in FP64 (the time-of-impact path
    is FP64 in both)."""
    def run(prec, continuous):
        b = continuous_batch(wall_scene()) if continuous else builder.scene_batch(wall_scene())
        if prec == "oracle":
            o = O.Oracle(b)
            step, state = (lambda: o.step(DT, 8, 3)), o.state
        else:
            e = N.Engine(prec)
            e.upload(b)
            step, state = (lambda: e.step(DT, 8, 3)), e.download
        for _ in range(40):
            step()
        return tuple(bool(x) for x in state().body_d[1:, L.BD_C_X] > 10.2)
    passed = run("oracle", False)
    assert all(passed) and run("fp64", False) == passed
    stopped = run("oracle", True)
    assert 0 < passed.count(True) - stopped.count(True) < 8
    assert run("fp64", True) == stopped
    assert run("fp32", True) == stopped


def test_drop_in_world_runs_the_reference_example_with_continuous_physics():
    """examples/hello.py's world, World((0, -10), True) with continuous physics on by default:
    a box falls onto a static ground box; the drop-in steps it on the device and matches the
    oracle."""
    world = b2.World((0, -10), True)
    assert world.continuous_physics
    ground = world.create_static_body(position=(0, -10))
    ground.create_polygon_fixture(box=(50, 10))
    body = world.create_dynamic_body(position=(0, 4))
    body.create_polygon_fixture(box=(1, 1), density=1, friction=0.3)
    o = O.Oracle(world.state())
    for _ in range(60):
        world.step(1.0 / 60, 6, 2)
        o.step(1.0 / 60, 6, 2)
    ref = o.state()
    assert abs(body.position.y - ref.body_d[1, L.BD_C_Y]) < 1e-9 and abs(body.angle - ref.body_d[1, L.BD_A]) < 1e-9
    assert abs(body.position.y - 1.0) < 0.02  # resting on the ground


def test_drop_in_world_listener_receives_the_sub_step_calls():
    """The projectiles scene on the drop-in World with every listener set: the calls made in the
    time-of-impact sub-steps (Contact.update of the impact contacts, Island.report) come after the
    step's other calls, as the oracle records them (phase 3, in call order)."""
    world = builder.build_world(dict(scenes.projectiles(), continuous=True), "fp64")
    fixtures = [f for bd in world.bodies for f in bd.fixtures]
    index = {id(f): k for k, f in enumerate(fixtures)}
    log = []

    def rec(kind):
        return lambda c, *rest: log.append((kind, index[id(c.fixture_a)], index[id(c.fixture_b)], c.touching,
                                            c.manifold.point_count))
    cm = world.contact_manager
    cm.begin_contact, cm.end_contact = rec(L.EV_BEGIN), rec(L.EV_END)
    cm.pre_solve, cm.post_solve = rec(L.EV_PRE_SOLVE), rec(L.EV_POST_SOLVE)
    o = O.Oracle(world.state())
    o.set_events(L.EV_ALL)
    sub_step_calls = 0
    for s in range(1, 61):
        world.step(DT, 8, 3)
        assert o.step(DT, 8, 3) == 0
        ei, _ = o.events(0)
        want = [(int(r[L.EI_KIND]), int(r[L.EI_FA]), int(r[L.EI_FB]), bool(r[L.EI_FLAGS] & L.CF_TOUCHING),
                 int(r[L.EI_MTYPE + 1])) for r in ei]
        assert log == want, s
        sub_step_calls += int(np.count_nonzero(ei[:, L.EI_PHASE] == 3))
        log.clear()
    assert sub_step_calls > 0
